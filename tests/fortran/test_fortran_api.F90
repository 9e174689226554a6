! Drives the library through the Fortran module: reads a .mmp file and an
! external-field file (one row x y z per MM atom, or 'none'), solves with the
! default solver, prints the polarization energy and the dipoles, then the
! potential and field of multipoles plus dipoles at two fixed external points
! (each evaluated twice into the same array, which must add up), then the
! gradient of the polarization energy (one row per MM atom).
program test_fortran_api
    use ommp_interface
    use iso_c_binding
    implicit none
    type(ommp_system), pointer :: s
    character(len=4096) :: mmp, eff
    real(ommp_real), allocatable :: ef(:,:), efmm(:,:), grd(:,:)
    real(ommp_real), pointer :: ipd(:,:,:)
    integer(ommp_integer), pointer :: pmm(:)
    integer :: i, k, npol, nmm, u
    real(ommp_real) :: ep, cext(3, 2), v(2), e(3, 2)

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
    ! the fixed-multipole energy first, as the reference's test_SI_potential.c:132 does
    write(*, '(A, ES24.15)') 'EM ', ommp_get_fixedelec_energy(s)
    call ommp_set_external_field(s, ef, OMMP_SOLVER_NONE, OMMP_MATV_NONE)
    ep = ommp_get_polelec_energy(s)
    write(*, '(A, ES24.15)') 'EP ', ep
    ipd => ommp_ipd(s)
    do k = 1, ommp_get_n_ipd(s)
        do i = 1, npol
            write(*, '(A, 3ES24.15)') 'IPD ', ipd(:, i, k)
        end do
    end do
    cext(:, 1) = [25.0_ommp_real, -31.0_ommp_real, 17.0_ommp_real]
    cext(:, 2) = [-40.0_ommp_real, 12.5_ommp_real, -22.0_ommp_real]
    v = 0.0_ommp_real
    e = 0.0_ommp_real
    call ommp_potential_mmpol2ext(s, 2, cext, v)
    call ommp_potential_mm2ext(s, 2, cext, v)
    call ommp_field_mmpol2ext(s, 2, cext, e)
    call ommp_field_pol2ext(s, 2, cext, e)
    do i = 1, 2
        write(*, '(A, 4ES24.15)') 'X2E ', v(i), e(:, i)
    end do
    allocate(grd(3, nmm))
    grd = 1.0_ommp_real
    call ommp_polelec_geomgrad(s, grd)
    do i = 1, nmm
        write(*, '(A, 3ES24.15)') 'GRD ', grd(:, i)
    end do
    call ommp_terminate(s)
end program
