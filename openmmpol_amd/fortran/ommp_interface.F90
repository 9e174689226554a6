! Fortran face of the MI355X polarization library.
!
! Procedures of the reference's Fortran module ommp_interface for the
! polarization path (src/mod_interface.F90:67-398 of the reference): same
! names, argument order and meaning.  Each one dispatches through
! ISO_C_BINDING to the C-ABI of libopenmmpol_amd.so (include/openmmpol.h),
! which runs the solve on the GPU.  The system is an opaque handle here
! instead of the reference's derived type; arrays keep the reference's
! column-major shapes: fields (3, pol_atoms), dipoles (3, pol_atoms, n_ipd),
! coordinates (3, mm_atoms).
module ommp_interface
    use iso_c_binding
    implicit none
    private

    integer, parameter, public :: ommp_integer = c_int32_t
    integer, parameter, public :: ommp_real = c_double
    integer, parameter, public :: ommp_logical = c_bool

    ! config/openmmpol_const.h.in:16-28 of the reference
    integer(ommp_integer), parameter, public :: OMMP_SOLVER_NONE = 0, OMMP_SOLVER_CG = 1, &
        OMMP_SOLVER_DIIS = 2, OMMP_SOLVER_INVERSION = 3, OMMP_SOLVER_DEFAULT = 1
    integer(ommp_integer), parameter, public :: OMMP_MATV_NONE = 0, OMMP_MATV_INCORE = 1, &
        OMMP_MATV_DIRECT = 2, OMMP_MATV_DEFAULT = 2
    integer(ommp_integer), parameter, public :: OMMP_VERBOSE_NONE = 0, OMMP_VERBOSE_LOW = 1, &
        OMMP_VERBOSE_HIGH = 2, OMMP_VERBOSE_DEBUG = 3
    integer(ommp_integer), parameter, public :: OMMP_AMOEBA_D = 1, OMMP_AMOEBA_P = 2

    type, public :: ommp_system
        type(c_ptr) :: h = c_null_ptr
    end type ommp_system

    interface
        function c_ommp_init_mmp(fname) bind(c, name='ommp_init_mmp') result(p)
            import :: c_ptr, c_char
            character(kind=c_char), dimension(*), intent(in) :: fname
            type(c_ptr) :: p
        end function
        subroutine c_ommp_terminate(s) bind(c, name='ommp_terminate')
            import :: c_ptr
            type(c_ptr), value :: s
        end subroutine
        subroutine c_ommp_set_verbose(v) bind(c, name='ommp_set_verbose')
            import :: c_int32_t
            integer(c_int32_t), value :: v
        end subroutine
        subroutine c_ommp_set_default_solver(s, v) bind(c, name='ommp_set_default_solver')
            import :: c_ptr, c_int32_t
            type(c_ptr), value :: s
            integer(c_int32_t), value :: v
        end subroutine
        subroutine c_ommp_set_default_matv(s, v) bind(c, name='ommp_set_default_matv')
            import :: c_ptr, c_int32_t
            type(c_ptr), value :: s
            integer(c_int32_t), value :: v
        end subroutine
        subroutine c_ommp_set_external_field(s, ef, solver, matv) bind(c, name='ommp_set_external_field')
            import :: c_ptr, c_double, c_int32_t
            type(c_ptr), value :: s
            real(c_double), intent(in) :: ef(*)
            integer(c_int32_t), value :: solver, matv
        end subroutine
        subroutine c_ommp_set_external_field_nomm(s, ef, solver, matv) &
                bind(c, name='ommp_set_external_field_nomm')
            import :: c_ptr, c_double, c_int32_t
            type(c_ptr), value :: s
            real(c_double), intent(in) :: ef(*)
            integer(c_int32_t), value :: solver, matv
        end subroutine
        function c_ommp_get_polelec_energy(s) bind(c, name='ommp_get_polelec_energy') result(e)
            import :: c_ptr, c_double
            type(c_ptr), value :: s
            real(c_double) :: e
        end function
        function c_ommp_get_fixedelec_energy(s) bind(c, name='ommp_get_fixedelec_energy') result(e)
            import :: c_ptr, c_double
            type(c_ptr), value :: s
            real(c_double) :: e
        end function
        function c_ommp_get_full_ele_energy(s) bind(c, name='ommp_get_full_ele_energy') result(e)
            import :: c_ptr, c_double
            type(c_ptr), value :: s
            real(c_double) :: e
        end function
        function c_ommp_get_pol_atoms(s) bind(c, name='ommp_get_pol_atoms') result(n)
            import :: c_ptr, c_int32_t
            type(c_ptr), value :: s
            integer(c_int32_t) :: n
        end function
        function c_ommp_get_mm_atoms(s) bind(c, name='ommp_get_mm_atoms') result(n)
            import :: c_ptr, c_int32_t
            type(c_ptr), value :: s
            integer(c_int32_t) :: n
        end function
        function c_ommp_get_n_ipd(s) bind(c, name='ommp_get_n_ipd') result(n)
            import :: c_ptr, c_int32_t
            type(c_ptr), value :: s
            integer(c_int32_t) :: n
        end function
        function c_ommp_get_ipd(s) bind(c, name='ommp_get_ipd') result(p)
            import :: c_ptr
            type(c_ptr), value :: s
            type(c_ptr) :: p
        end function
        function c_ommp_get_polar_mm(s) bind(c, name='ommp_get_polar_mm') result(p)
            import :: c_ptr
            type(c_ptr), value :: s
            type(c_ptr) :: p
        end function
        function c_ommp_get_zmm(s) bind(c, name='ommp_get_zmm') result(p)
            import :: c_ptr
            type(c_ptr), value :: s
            type(c_ptr) :: p
        end function
        subroutine c_ommp_update_coordinates(s, c) bind(c, name='ommp_update_coordinates')
            import :: c_ptr, c_double
            type(c_ptr), value :: s
            real(c_double), intent(in) :: c(*)
        end subroutine
        subroutine c_ommp_turn_pol_off(s, n, nopol) bind(c, name='ommp_turn_pol_off')
            import :: c_ptr, c_int32_t
            type(c_ptr), value :: s
            integer(c_int32_t), value :: n
            integer(c_int32_t), intent(in) :: nopol(*)
        end subroutine
        subroutine c_ommp_potential_mmpol2ext(s, n, cext, v) bind(c, name='ommp_potential_mmpol2ext')
            import :: c_ptr, c_double, c_int32_t
            type(c_ptr), value :: s
            integer(c_int32_t), value :: n
            real(c_double), intent(in) :: cext(*)
            real(c_double), intent(inout) :: v(*)
        end subroutine
        subroutine c_ommp_potential_mm2ext(s, n, cext, v) bind(c, name='ommp_potential_mm2ext')
            import :: c_ptr, c_double, c_int32_t
            type(c_ptr), value :: s
            integer(c_int32_t), value :: n
            real(c_double), intent(in) :: cext(*)
            real(c_double), intent(inout) :: v(*)
        end subroutine
        subroutine c_ommp_potential_pol2ext(s, n, cext, v) bind(c, name='ommp_potential_pol2ext')
            import :: c_ptr, c_double, c_int32_t
            type(c_ptr), value :: s
            integer(c_int32_t), value :: n
            real(c_double), intent(in) :: cext(*)
            real(c_double), intent(inout) :: v(*)
        end subroutine
        subroutine c_ommp_field_mmpol2ext(s, n, cext, e) bind(c, name='ommp_field_mmpol2ext')
            import :: c_ptr, c_double, c_int32_t
            type(c_ptr), value :: s
            integer(c_int32_t), value :: n
            real(c_double), intent(in) :: cext(*)
            real(c_double), intent(inout) :: e(*)
        end subroutine
        subroutine c_ommp_field_mm2ext(s, n, cext, e) bind(c, name='ommp_field_mm2ext')
            import :: c_ptr, c_double, c_int32_t
            type(c_ptr), value :: s
            integer(c_int32_t), value :: n
            real(c_double), intent(in) :: cext(*)
            real(c_double), intent(inout) :: e(*)
        end subroutine
        subroutine c_ommp_field_pol2ext(s, n, cext, e) bind(c, name='ommp_field_pol2ext')
            import :: c_ptr, c_double, c_int32_t
            type(c_ptr), value :: s
            integer(c_int32_t), value :: n
            real(c_double), intent(in) :: cext(*)
            real(c_double), intent(inout) :: e(*)
        end subroutine
        subroutine c_ommp_polelec_geomgrad(s, grd) bind(c, name='ommp_polelec_geomgrad')
            import :: c_ptr, c_double
            type(c_ptr), value :: s
            real(c_double), intent(out) :: grd(*)
        end subroutine
        subroutine c_ommp_rotation_geomgrad(s, e, egrd, grd) bind(c, name='ommp_rotation_geomgrad')
            import :: c_ptr, c_double
            type(c_ptr), value :: s
            real(c_double), intent(in) :: e(*), egrd(*)
            real(c_double), intent(out) :: grd(*)
        end subroutine
    end interface

    public :: ommp_potential_mmpol2ext, ommp_potential_mm2ext, ommp_potential_pol2ext
    public :: ommp_field_mmpol2ext, ommp_field_mm2ext, ommp_field_pol2ext
    public :: ommp_polelec_geomgrad, ommp_rotation_geomgrad
    public :: ommp_init_mmp, ommp_terminate, ommp_set_verbose
    public :: ommp_set_default_solver, ommp_set_default_matv
    public :: ommp_set_external_field, ommp_set_external_field_nomm
    public :: ommp_get_polelec_energy, ommp_get_fixedelec_energy, ommp_get_full_ele_energy
    public :: ommp_get_pol_atoms, ommp_get_mm_atoms, ommp_get_n_ipd
    public :: ommp_ipd, ommp_polar_mm, ommp_zmm, ommp_update_coordinates, ommp_turn_pol_off

contains

    ! mod_interface.F90:89-100
    subroutine ommp_init_mmp(s, filename)
        type(ommp_system), pointer, intent(inout) :: s
        character(len=*), intent(in) :: filename
        allocate(s)
        s%h = c_ommp_init_mmp(trim(filename)//c_null_char)
    end subroutine

    ! mod_interface.F90:161-170
    subroutine ommp_terminate(s)
        type(ommp_system), pointer, intent(inout) :: s
        call c_ommp_terminate(s%h)
        deallocate(s)
    end subroutine

    subroutine ommp_set_verbose(v)
        integer(ommp_integer), intent(in) :: v
        call c_ommp_set_verbose(v)
    end subroutine

    subroutine ommp_set_default_solver(s, solver)
        type(ommp_system), intent(in) :: s
        integer(ommp_integer), intent(in), value :: solver
        call c_ommp_set_default_solver(s%h, solver)
    end subroutine

    subroutine ommp_set_default_matv(s, matv)
        type(ommp_system), intent(in) :: s
        integer(ommp_integer), intent(in), value :: matv
        call c_ommp_set_default_matv(s%h, matv)
    end subroutine

    ! mod_interface.F90:184-223: add_mm_field (default .true.) selects the
    ! MM-field variant; .false. is ommp_set_external_field_nomm.
    subroutine ommp_set_external_field(s, ext_field, solver, matv, add_mm_field)
        type(ommp_system), intent(in) :: s
        real(ommp_real), intent(in) :: ext_field(:,:)
        integer(ommp_integer), intent(in), value :: solver, matv
        logical, intent(in), optional :: add_mm_field
        logical :: mm
        mm = .true.
        if(present(add_mm_field)) mm = add_mm_field
        if(mm) then
            call c_ommp_set_external_field(s%h, ext_field, solver, matv)
        else
            call c_ommp_set_external_field_nomm(s%h, ext_field, solver, matv)
        end if
    end subroutine

    subroutine ommp_set_external_field_nomm(s, ext_field, solver, matv)
        type(ommp_system), intent(in) :: s
        real(ommp_real), intent(in) :: ext_field(:,:)
        integer(ommp_integer), intent(in), value :: solver, matv
        call c_ommp_set_external_field_nomm(s%h, ext_field, solver, matv)
    end subroutine

    ! mod_interface.F90:343-370
    function ommp_get_polelec_energy(s) result(e)
        type(ommp_system), intent(in) :: s
        real(ommp_real) :: e
        e = c_ommp_get_polelec_energy(s%h)
    end function

    ! mod_interface.F90:372-383
    function ommp_get_fixedelec_energy(s) result(e)
        type(ommp_system), intent(in) :: s
        real(ommp_real) :: e
        e = c_ommp_get_fixedelec_energy(s%h)
    end function

    ! mod_interface.F90:388-398
    function ommp_get_full_ele_energy(s) result(e)
        type(ommp_system), intent(in) :: s
        real(ommp_real) :: e
        e = c_ommp_get_full_ele_energy(s%h)
    end function

    function ommp_get_pol_atoms(s) result(n)
        type(ommp_system), intent(in) :: s
        integer(ommp_integer) :: n
        n = c_ommp_get_pol_atoms(s%h)
    end function

    function ommp_get_mm_atoms(s) result(n)
        type(ommp_system), intent(in) :: s
        integer(ommp_integer) :: n
        n = c_ommp_get_mm_atoms(s%h)
    end function

    function ommp_get_n_ipd(s) result(n)
        type(ommp_system), intent(in) :: s
        integer(ommp_integer) :: n
        n = c_ommp_get_n_ipd(s%h)
    end function

    ! The reference exposes eel%ipd(3, pol_atoms, n_ipd); here a pointer view of
    ! the library-owned host buffer (refreshed after every solve).
    function ommp_ipd(s) result(p)
        type(ommp_system), intent(in) :: s
        real(ommp_real), pointer :: p(:,:,:)
        call c_f_pointer(c_ommp_get_ipd(s%h), p, [3, ommp_get_pol_atoms(s), ommp_get_n_ipd(s)])
    end function

    ! 0-based MM index of each polarizable site (C_ommp_get_polar_mm)
    function ommp_polar_mm(s) result(p)
        type(ommp_system), intent(in) :: s
        integer(ommp_integer), pointer :: p(:)
        call c_f_pointer(c_ommp_get_polar_mm(s%h), p, [ommp_get_pol_atoms(s)])
    end function

    ! mod_c_interface.F90:950-959: atomic numbers [mm_atoms] (zero from .mmp input)
    function ommp_zmm(s) result(p)
        type(ommp_system), intent(in) :: s
        integer(ommp_integer), pointer :: p(:)
        call c_f_pointer(c_ommp_get_zmm(s%h), p, [ommp_get_mm_atoms(s)])
    end function

    ! mod_interface.F90:134-168: the listed MM atoms (1-based) lose their
    ! polarizability (see include/openmmpol.h)
    subroutine ommp_turn_pol_off(s, n, nopol)
        type(ommp_system), intent(in) :: s
        integer(ommp_integer), intent(in) :: n
        integer(ommp_integer), intent(in) :: nopol(n)
        call c_ommp_turn_pol_off(s%h, n, nopol)
    end subroutine

    ! mod_mmpol.F90:325-372
    subroutine ommp_update_coordinates(s, new_c)
        type(ommp_system), intent(in) :: s
        real(ommp_real), intent(in) :: new_c(:,:)
        call c_ommp_update_coordinates(s%h, new_c)
    end subroutine

    ! mod_interface.F90:250-295: results are added to v
    subroutine ommp_potential_mmpol2ext(s, n, cext, v)
        type(ommp_system), intent(in) :: s
        integer(ommp_integer), intent(in) :: n
        real(ommp_real), intent(in) :: cext(3, n)
        real(ommp_real), intent(inout) :: v(n)
        call c_ommp_potential_mmpol2ext(s%h, n, cext, v)
    end subroutine

    ! mod_interface.F90:250-295: results are added to v
    subroutine ommp_potential_mm2ext(s, n, cext, v)
        type(ommp_system), intent(in) :: s
        integer(ommp_integer), intent(in) :: n
        real(ommp_real), intent(in) :: cext(3, n)
        real(ommp_real), intent(inout) :: v(n)
        call c_ommp_potential_mm2ext(s%h, n, cext, v)
    end subroutine

    ! mod_interface.F90:250-295: results are added to v
    subroutine ommp_potential_pol2ext(s, n, cext, v)
        type(ommp_system), intent(in) :: s
        integer(ommp_integer), intent(in) :: n
        real(ommp_real), intent(in) :: cext(3, n)
        real(ommp_real), intent(inout) :: v(n)
        call c_ommp_potential_pol2ext(s%h, n, cext, v)
    end subroutine

    ! mod_interface.F90:297-341: results are added to e
    subroutine ommp_field_mmpol2ext(s, n, cext, e)
        type(ommp_system), intent(in) :: s
        integer(ommp_integer), intent(in), value :: n
        real(ommp_real), intent(in) :: cext(3, n)
        real(ommp_real), intent(inout) :: e(3, n)
        call c_ommp_field_mmpol2ext(s%h, n, cext, e)
    end subroutine

    ! mod_interface.F90:297-341: results are added to e
    subroutine ommp_field_mm2ext(s, n, cext, e)
        type(ommp_system), intent(in) :: s
        integer(ommp_integer), intent(in), value :: n
        real(ommp_real), intent(in) :: cext(3, n)
        real(ommp_real), intent(inout) :: e(3, n)
        call c_ommp_field_mm2ext(s%h, n, cext, e)
    end subroutine

    ! mod_interface.F90:297-341: results are added to e
    subroutine ommp_field_pol2ext(s, n, cext, e)
        type(ommp_system), intent(in) :: s
        integer(ommp_integer), intent(in), value :: n
        real(ommp_real), intent(in) :: cext(3, n)
        real(ommp_real), intent(inout) :: e(3, n)
        call c_ommp_field_pol2ext(s%h, n, cext, e)
    end subroutine

    ! mod_interface.F90:671-683: grd(3, mm_atoms) = 0, then polelec_geomgrad
    ! (Hartree/bohr)
    subroutine ommp_polelec_geomgrad(s, grd)
        type(ommp_system), intent(in) :: s
        real(ommp_real), intent(out), contiguous :: grd(:, :)
        call c_ommp_polelec_geomgrad(s%h, grd)
    end subroutine

    ! mod_interface.F90:699-710: grd = 0, then rotation_geomgrad for the field
    ! e(3, mm_atoms) and field gradient egrd(6, mm_atoms)
    subroutine ommp_rotation_geomgrad(s, e, egrd, grd)
        type(ommp_system), intent(in) :: s
        real(ommp_real), intent(in), contiguous :: e(:, :), egrd(:, :)
        real(ommp_real), intent(out), contiguous :: grd(:, :)
        call c_ommp_rotation_geomgrad(s%h, e, egrd, grd)
    end subroutine

end module ommp_interface
