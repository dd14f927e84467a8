!=======================================================================
! Host-only check of the Fortran binding: fills godar_params through the
! library (godar_params_default needs no GPU) and prints every field, so a
! test can compare the Fortran view of the struct with the C one.
!=======================================================================
program abi_check
    use iso_c_binding
    use godar_gpu
    implicit none
    type(godar_params) :: p
    integer(c_int) :: rc
    rc = godar_params_default(p)
    write(*, '(A, I0)') 'rc ', rc
    write(*, '(A, I0)') 'sizeof ', c_sizeof(p)
    write(*, '(A, 33(1X, ES24.17))') 'doubles', p%dt, p%rtree, p%nx, p%ny, p%rhoair, p%rhoice, &
        p%rhowater, p%Cdair, p%Cdwater, p%Csair, p%Cswater, p%z0w, p%lat, p%e_modul, &
        p%poiss_ratio, p%friction_coeff, p%rest_coeff, p%sigmanc_crit, p%eb, p%lambda_rb, &
        p%lambda_lb, p%sigmatb_crit, p%sigmacb_crit, p%tau_crit, p%bond_lim, p%dmax, p%dtd, &
        p%dth, p%ua, p%va, p%uw, p%vw, p%bond_frac
    write(*, '(A, 6(1X, I0), 1X, I0)') 'ints', p%ntree, p%slipping, p%cohesion, p%ridging, &
        p%shelter, p%restart, p%bond_seed
end program abi_check
