"""Copy the reference's own hot-path test fixtures into tests/golden/reference/.

These are data files the reference's ctest suite holds (inputs, external
fields, golden induced dipoles and energies, prepared-state summaries, and the
FULL_POTENTIAL.ref energy/gradient listings of the AMOEBA inputs; the 1UBQ one
is what the reference's 1UBQ_AMOEBA_XYZ_geomgrad test checks its fast multipole
gradient against), stored gzipped.  Run once in a container that has the reference tree mounted:

    python tests/golden/collect_reference_fixtures.py /root/reference

Source rows: tests/test_list:3-73 of the reference.
"""
import gzip
import os
import shutil
import sys

FILES = {
    "N-methylacetamide": [
        "input_AMOEBA.mmp", "input_AMOEBA_CUT.mmp", "input_WANG_AL.mmp", "input_WANG_AL_CUT.mmp",
        "EF_1.txt",
        "IPD_0_AMOEBA.ref", "IPD_1_AMOEBA.ref", "IPD_0_AMOEBA_CUT.ref", "IPD_1_AMOEBA_CUT.ref",
        "IPD_0_WANG_AL.ref", "IPD_1_WANG_AL.ref", "IPD_0_WANG_AL_CUT.ref", "IPD_1_WANG_AL_CUT.ref",
        "ENE_0_AMOEBA.ref", "ENE_1_AMOEBA.ref", "ENE_0_WANG_AL.ref", "ENE_1_WANG_AL.ref",
        "summary_AMOEBA.ref", "summary_WANG_AL.ref",
        "FULL_POTENTIAL.ref",
    ],
    "1crn": [
        "input_AMOEBA.mmp", "input_WANG_AL.mmp", "EF_1.txt",
        "IPD_0_AMOEBA.ref", "IPD_1_AMOEBA.ref", "IPD_0_WANG_AL.ref", "IPD_1_WANG_AL.ref",
        "ENE_0_AMOEBA.ref", "ENE_1_AMOEBA.ref", "ENE_0_WANG_AL.ref", "ENE_1_WANG_AL.ref",
        "summary_AMOEBA.ref", "summary_WANG_AL.ref",
        "FULL_POTENTIAL.ref",
    ],
    "1ubq": [
        "input_AMOEBA.mmp", "input_WANG_AL.mmp", "EF_1.txt",
        "IPD_0_AMOEBA.ref", "IPD_1_AMOEBA.ref", "IPD_0_WANG_AL.ref", "IPD_1_WANG_AL.ref",
        "ENE_0_AMOEBA.ref", "ENE_1_AMOEBA.ref", "ENE_0_WANG_AL.ref", "ENE_1_WANG_AL.ref",
        "FULL_POTENTIAL.ref",
    ],
    # 18,331 MM sites, 37 polarizable: the reference's only large case with many
    # non-polarizable sources (tests/test_list:59-73; Wang-AL only, the AMOEBA
    # CUT10 input is built from Tinker files this tree cannot load)
    "1ao6": [
        "input_WANG_AL_CUT10.mmp", "EF_1.txt",
        "IPD_0_WANG_AL_CUT10.ref", "IPD_1_WANG_AL_CUT10.ref",
        "ENE_0_WANG_AL_CUT10.ref", "ENE_1_WANG_AL_CUT10.ref",
    ],
}


def main(ref_root):
    out_root = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference")
    for system, names in FILES.items():
        os.makedirs(os.path.join(out_root, system), exist_ok=True)
        for name in names:
            src = os.path.join(ref_root, "tests", system, name)
            dst = os.path.join(out_root, system, name + ".gz")
            with open(src, "rb") as fi, gzip.GzipFile(dst, "wb", mtime=0) as fo:
                shutil.copyfileobj(fi, fo)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
