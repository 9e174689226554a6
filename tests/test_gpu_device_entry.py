"""The device-pointer entries, which take fields already in device memory (the
benchmark's steps run through them), give the results of the host-pointer
entries: ommp_amd_set_external_field_device against ommp_set_external_field,
and ommp_amd_set_external_fields_multi_device (dipoles left in device memory)
against ommp_amd_set_external_fields_multi, on a 3,000-site sphere in both
precisions; ommp_amd_device_ipd points at the dipoles ommp_get_ipd returns.
Bar: the same solver on the same data, so 1e-12 of the largest dipole
component (the single-field solves repeat bit for bit)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _Device:
    """Plain device allocations through the HIP runtime."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.ptrs = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) == 0
        self.ptrs.append(p)
        return p.value

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = self.alloc(a.nbytes)
        assert self.hip.hipMemcpy(ctypes.c_void_p(p), a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes),
                                  1) == 0
        return p

    def download(self, p, shape):
        out = np.empty(shape)
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(p),
                                  ctypes.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def _sphere(seed):
    from openmmpol_amd import synth
    n = 3000
    spec = synth.sphere_system(n, (3.0 * n / (4.0 * np.pi * 0.0996)) ** (1.0 / 3.0), seed=seed)
    return synth.make_system(spec)


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_single_field_device_entry_matches_host_entry(mode):
    s = _sphere(31)
    s.set_precision(mode)
    ext = np.random.default_rng(3).normal(scale=1e-3, size=(s.pol_atoms, 3))
    dev = _Device()
    try:
        s.reset_guess(True)
        s.set_external_field(ext)
        e_host = s.get_polelec_energy()
        ipd_host = np.array(s.ipd)
        s.reset_guess(True)
        s.set_external_field_device(dev.upload(ext))
        e_dev = s.get_polelec_energy()
        # the dipoles read where the solver left them in device memory
        ipd_devmem = dev.download(s.device_ipd_ptr(), (s.n_ipd, s.pol_atoms, 3))
        ipd_dev = np.array(s.ipd)
    finally:
        dev.free()
    assert np.array_equal(ipd_devmem, ipd_dev)
    diff = np.abs(ipd_dev - ipd_host).max() / np.abs(ipd_host).max()
    print(f"{mode}: dipoles {diff:.2e} of the largest, bitwise {np.array_equal(ipd_dev, ipd_host)}, "
          f"E_pol {e_host:.12e} / {e_dev:.12e}", flush=True)
    assert diff <= 1e-12
    assert abs(e_dev - e_host) <= 1e-12 * abs(e_host)
    s.terminate()


def test_multi_field_device_entry_matches_host_entry():
    s = _sphere(32)
    nrhs = 4
    ext = np.random.default_rng(4).normal(scale=1e-3, size=(nrhs, s.pol_atoms, 3))
    s.set_external_field(np.zeros((s.pol_atoms, 3)))  # a common starting point for both calls
    ipd_host, ep_host = s.set_external_fields_multi(ext)
    dev = _Device()
    try:
        shape = (nrhs, s.n_ipd, s.pol_atoms, 3)
        d_ipd = dev.alloc(8 * int(np.prod(shape)))
        ep_dev = s.set_external_fields_multi_device(nrhs, dev.upload(ext), d_ipd)
        ipd_dev = dev.download(d_ipd, shape)
    finally:
        dev.free()
    diff = np.abs(ipd_dev - ipd_host).max() / np.abs(ipd_host).max()
    print(f"multi: dipoles {diff:.2e} of the largest, bitwise {np.array_equal(ipd_dev, ipd_host)}", flush=True)
    assert diff <= 1e-12
    assert np.abs(ep_dev - ep_host).max() <= 1e-12 * np.abs(ep_host).max()
    s.terminate()
