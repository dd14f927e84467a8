!=======================================================================
! iso_c_binding interface of libgodar_amd.so (include/godar_gpu.h) for the
! GODAR Fortran driver.
!
! The reference has no FFI: its time loop calls one external subroutine,
!     call stepper (tstep, restart)                     model/src/godar.f90:193
! which works on the COMMON blocks of model/inc/CB_*.h.  This module binds
! the C ABI that replaces it; stepper_gpu.f90 is the drop-in subroutine with
! the same arguments built on it.
!
! Indices: the C ABI is 0-based (member-local floe indices); the helpers at
! the end of this module take and return 1-based Fortran indices.
!=======================================================================
module godar_gpu
    use iso_c_binding
    implicit none
    private :: c_strlen

    ! status codes
    integer(c_int), parameter :: GODAR_OK = 0, GODAR_ERR_ARG = 1, GODAR_ERR_HIP = 2, &
                                 GODAR_ERR_CAPACITY = 3, GODAR_ERR_NOMEM = 4, GODAR_ERR_STATE = 5
    ! arithmetic of the force laws
    integer(c_int), parameter :: GODAR_FP64 = 0, GODAR_MIXED = 1
    ! pair kinds
    integer(c_int), parameter :: GODAR_PAIRS_CANDIDATE = 0, GODAR_PAIRS_CONTACT = 1, &
                                 GODAR_PAIRS_BOND = 2
    ! columns of godar_get_forces (1-based rows of the (33, n) block as seen
    ! from Fortran, i.e. forces(k, i) with k below)
    integer, parameter :: GF_TFX = 1, GF_TFY = 2, GF_M = 3, GF_FCX = 4, GF_FCY = 5, GF_MC = 6, &
                          GF_FBX = 7, GF_FBY = 8, GF_MB = 9, GF_FAX = 10, GF_FAY = 11, GF_MA = 12, &
                          GF_FWX = 13, GF_FWY = 14, GF_MW = 15, GF_FCORX = 16, GF_FCORY = 17, &
                          GF_FX_BC = 18, GF_FY_BC = 19, GF_M_BC = 20, GF_HSFA_MIN = 21, &
                          GF_HSFW_MIN = 22, GF_TP = 23, GF_TSIGXX = 24, GF_TSIGYY = 25, &
                          GF_TSIGXY = 26, GF_TSIGYX = 27, GF_TAC = 28, GF_PC = 29, GF_TAB = 30, &
                          GF_PB = 31, GF_TA_BC = 32, GF_P_BC = 33
    integer, parameter :: GODAR_NFORCES = 33, GODAR_NBONDSTATE = 11

    ! settings of stepper_gpu (stepper_gpu.f90), read at its first call:
    ! arithmetic of the force laws, and whether the COMMON blocks are
    ! refreshed after every step instead of on output steps only
    integer(c_int32_t) :: godar_stepper_precision = GODAR_FP64
    logical :: godar_sync_every = .false.

    ! godar_params of include/godar_gpu.h, field for field
    type, bind(C) :: godar_params
        real(c_double) :: dt, rtree, nx, ny
        real(c_double) :: rhoair, rhoice, rhowater
        real(c_double) :: Cdair, Cdwater, Csair, Cswater
        real(c_double) :: z0w, lat
        real(c_double) :: e_modul, poiss_ratio, friction_coeff, rest_coeff, sigmanc_crit
        real(c_double) :: eb, lambda_rb, lambda_lb, sigmatb_crit, sigmacb_crit, tau_crit
        real(c_double) :: bond_lim, dmax, dtd, dth
        real(c_double) :: ua, va, uw, vw
        real(c_double) :: bond_frac
        integer(c_int32_t) :: ntree, slipping, cohesion, ridging, shelter, restart
        integer(c_int64_t) :: bond_seed          ! uint64_t in C
    end type godar_params

    interface
        integer(c_int) function godar_params_default(p) bind(C, name="godar_params_default")
            import :: c_int, godar_params
            type(godar_params), intent(out) :: p
        end function

        integer(c_int) function godar_create(max_floes, max_members, precision, ctx) &
                bind(C, name="godar_create")
            import :: c_int, c_int32_t, c_int64_t, c_ptr
            integer(c_int64_t), value :: max_floes
            integer(c_int32_t), value :: max_members, precision
            type(c_ptr), intent(out) :: ctx
        end function

        integer(c_int) function godar_destroy(ctx) bind(C, name="godar_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function

        type(c_ptr) function godar_last_error_c(ctx) bind(C, name="godar_last_error")
            import :: c_ptr
            type(c_ptr), value :: ctx
        end function

        integer(c_int) function godar_member_add(ctx, p, n, x, y, u, v, theta, omega, r, h, &
                                                 tstep0, member) bind(C, name="godar_member_add")
            import :: c_int, c_int32_t, c_double, c_ptr, godar_params
            type(c_ptr), value :: ctx
            type(godar_params), intent(in) :: p
            integer(c_int32_t), value :: n, tstep0
            real(c_double), intent(in) :: x(*), y(*), u(*), v(*), theta(*), omega(*), r(*), h(*)
            integer(c_int32_t), intent(out) :: member
        end function

        integer(c_int) function godar_member_remove(ctx, member) bind(C, name="godar_member_remove")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: member
        end function

        integer(c_int) function godar_member_footprint(ctx, n, bytes) &
                bind(C, name="godar_member_footprint")
            import :: c_int, c_int32_t, c_int64_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: n
            integer(c_int64_t), intent(out) :: bytes
        end function

        integer(c_int) function godar_capacity_left(ctx, floes, members) &
                bind(C, name="godar_capacity_left")
            import :: c_int, c_int32_t, c_int64_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int64_t), intent(out) :: floes
            integer(c_int32_t), intent(out) :: members
        end function

        integer(c_int) function godar_set_forcing(ctx, member, ua, va, uw, vw) &
                bind(C, name="godar_set_forcing")
            import :: c_int, c_int32_t, c_double, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: member
            real(c_double), value :: ua, va, uw, vw
        end function

        integer(c_int) function godar_set_wind_grid(ctx, member, nxn, nyn, x0, y0, dx, dy, ua, va) &
                bind(C, name="godar_set_wind_grid")
            import :: c_int, c_int32_t, c_double, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: member, nxn, nyn
            real(c_double), value :: x0, y0, dx, dy
            real(c_double), intent(in) :: ua(*), va(*)
        end function

        integer(c_int) function godar_set_bonds_c(ctx, member, nb, bi, bj, state) &
                bind(C, name="godar_set_bonds")
            import :: c_int, c_int32_t, c_double, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: member, nb
            integer(c_int32_t), intent(in) :: bi(*), bj(*)
            real(c_double), intent(in) :: state(*)
        end function

        integer(c_int) function godar_step(ctx, nsteps) bind(C, name="godar_step")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: nsteps
        end function

        integer(c_int) function godar_step_async(ctx, nsteps) bind(C, name="godar_step_async")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: nsteps
        end function

        integer(c_int) function godar_sync(ctx) bind(C, name="godar_sync")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function

        integer(c_int) function godar_member_tstep(ctx, member, tstep) &
                bind(C, name="godar_member_tstep")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: member
            integer(c_int32_t), intent(out) :: tstep
        end function

        integer(c_int) function godar_member_run_until(ctx, member, tlast) &
                bind(C, name="godar_member_run_until")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: member, tlast
        end function

        integer(c_int) function godar_member_tsteps(ctx, tsteps) bind(C, name="godar_member_tsteps")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), intent(out) :: tsteps(*)
        end function

        integer(c_int) function godar_reserve(ctx, outer, inner, hist, bonds) &
                bind(C, name="godar_reserve")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: outer, inner, hist, bonds
        end function

        integer(c_int) function godar_get_floes(ctx, member, x, y, u, v, theta, omega, r, h, &
                                                mass, inertia, hfa, hfw) &
                bind(C, name="godar_get_floes")
            import :: c_int, c_int32_t, c_double, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: member
            real(c_double), intent(out) :: x(*), y(*), u(*), v(*), theta(*), omega(*), r(*), &
                                           h(*), mass(*), inertia(*), hfa(*), hfw(*)
        end function

        integer(c_int) function godar_record_forces(ctx, on) bind(C, name="godar_record_forces")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: on
        end function

        ! f: the C (33, n) row-major block = a Fortran (n, 33) array
        integer(c_int) function godar_get_forces(ctx, member, f) bind(C, name="godar_get_forces")
            import :: c_int, c_int32_t, c_double, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: member
            real(c_double), intent(out) :: f(*)
        end function

        integer(c_int) function godar_get_walls(ctx, member, w) bind(C, name="godar_get_walls")
            import :: c_int, c_int32_t, c_double, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: member
            real(c_double), intent(out) :: w(*)
        end function

        integer(c_int) function godar_get_pairs_c(ctx, member, kind, cap, pi, pj, count) &
                bind(C, name="godar_get_pairs")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: member, kind, cap
            integer(c_int32_t), intent(out) :: pi(*), pj(*)
            integer(c_int32_t), intent(out) :: count
        end function

        integer(c_int) function godar_get_bond_state(ctx, member, cap, state, count) &
                bind(C, name="godar_get_bond_state")
            import :: c_int, c_int32_t, c_double, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: member, cap
            real(c_double), intent(out) :: state(*)
            integer(c_int32_t), intent(out) :: count
        end function

        integer(c_int) function godar_get_stats(ctx, member, stats) bind(C, name="godar_get_stats")
            import :: c_int, c_int32_t, c_int64_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: member
            integer(c_int64_t), intent(out) :: stats(4)
        end function

        integer(c_int) function godar_set_block(ctx, threads) bind(C, name="godar_set_block")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: threads
        end function

        integer(c_int) function godar_use_graphs(ctx, on) bind(C, name="godar_use_graphs")
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: on
        end function

        function c_strlen(s) bind(C, name="strlen")
            import :: c_ptr, c_size_t
            type(c_ptr), value :: s
            integer(c_size_t) :: c_strlen
        end function
    end interface

contains

    ! Text of the context's last error.
    function godar_last_error(ctx) result(msg)
        type(c_ptr), intent(in) :: ctx
        character(len=:), allocatable :: msg
        type(c_ptr) :: p
        character(kind=c_char), pointer :: chars(:)
        integer :: k, l
        p = godar_last_error_c(ctx)
        if (.not. c_associated(p)) then
            msg = ''
            return
        end if
        l = int(c_strlen(p))
        call c_f_pointer(p, chars, [l])
        allocate(character(len=l) :: msg)
        do k = 1, l
            msg(k:k) = chars(k)
        end do
    end function

    ! Stop with the library's message when rc is not GODAR_OK.
    subroutine godar_check(rc, ctx, what)
        integer(c_int), intent(in) :: rc
        type(c_ptr), intent(in) :: ctx
        character(len=*), intent(in) :: what
        if (rc /= GODAR_OK) then
            write(*, '(A, A, A, I0, A, A)') 'godar: ', what, ' failed (status ', rc, '): ', &
                godar_last_error(ctx)
            stop 1
        end if
    end subroutine

    ! Bonds with 1-based floe indices i < j; state(11, nb) as in godar_set_bonds.
    integer(c_int) function godar_set_bonds(ctx, member, nb, bi, bj, state) result(rc)
        type(c_ptr), intent(in) :: ctx
        integer(c_int32_t), intent(in) :: member, nb
        integer, intent(in) :: bi(:), bj(:)
        real(c_double), intent(in) :: state(:, :)
        integer(c_int32_t), allocatable :: i0(:), j0(:)
        allocate(i0(max(nb, 1)), j0(max(nb, 1)))
        i0(1:nb) = int(bi(1:nb) - 1, c_int32_t)
        j0(1:nb) = int(bj(1:nb) - 1, c_int32_t)
        rc = godar_set_bonds_c(ctx, member, nb, i0, j0, state)
    end function

    ! Pairs of a kind with 1-based indices, sorted by (i, j); count gets the
    ! total, at most size(pi) pairs are written.
    integer(c_int) function godar_get_pairs(ctx, member, kind, pi, pj, count) result(rc)
        type(c_ptr), intent(in) :: ctx
        integer(c_int32_t), intent(in) :: member, kind
        integer(c_int32_t), intent(out) :: pi(:), pj(:)
        integer(c_int32_t), intent(out) :: count
        integer :: k
        rc = godar_get_pairs_c(ctx, member, kind, int(size(pi), c_int32_t), pi, pj, count)
        if (rc /= GODAR_OK) return
        do k = 1, min(int(count), size(pi))
            pi(k) = pi(k) + 1
            pj(k) = pj(k) + 1
        end do
    end function

end module godar_gpu
