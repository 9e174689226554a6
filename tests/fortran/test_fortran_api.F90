! Drives the library through the Fortran module: reads a .mmp file and an
! external-field file (one row x y z per MM atom, or 'none'), solves with the
! default solver, prints the polarization energy and the dipoles.
program test_fortran_api
    use ommp_interface
    use iso_c_binding
    implicit none
    type(ommp_system), pointer :: s
    character(len=4096) :: mmp, eff
    real(ommp_real), allocatable :: ef(:,:), efmm(:,:)
    real(ommp_real), pointer :: ipd(:,:,:)
    integer(ommp_integer), pointer :: pmm(:)
    integer :: i, k, npol, nmm, u
    real(ommp_real) :: ep

    call get_command_argument(1, mmp)
    call get_command_argument(2, eff)
    call ommp_set_verbose(OMMP_VERBOSE_NONE)
    call ommp_init_mmp(s, trim(mmp))
    npol = ommp_get_pol_atoms(s)
    nmm = ommp_get_mm_atoms(s)
    allocate(ef(3, npol))
    ef = 0.0_ommp_real
    if (trim(eff) /= 'none') then
        allocate(efmm(3, nmm))
        open(newunit=u, file=trim(eff), action='read')
        do i = 1, nmm
            read(u, *) efmm(:, i)
        end do
        close(u)
        pmm => ommp_polar_mm(s)
        do i = 1, npol
            ef(:, i) = efmm(:, pmm(i) + 1)
        end do
    end if
    call ommp_set_external_field(s, ef, OMMP_SOLVER_NONE, OMMP_MATV_NONE)
    ep = ommp_get_polelec_energy(s)
    write(*, '(A, ES24.15)') 'EP ', ep
    ipd => ommp_ipd(s)
    do k = 1, ommp_get_n_ipd(s)
        do i = 1, npol
            write(*, '(A, 3ES24.15)') 'IPD ', ipd(:, i, k)
        end do
    end do
    call ommp_terminate(s)
end program
