! ---------------------------------------------------------------------------
! TEST INFRASTRUCTURE ONLY.  Never linked into the product library.
!
! Driver around GODAR's own physics routines, which are compiled unmodified
! from the reference sources (model/src/{physics,contact_forces,bonds,ridging,
! forcing,boundaries,reset,integration,par_get}.f90) together with this file.
! The reference's step (model/src/stepper.f90) cannot be compiled here: it
! needs the coretran k-d tree and an MPI mpi_f08 module, neither of which is
! in the image.  This file therefore restates only stepper.f90's control flow
! for one MPI rank and one OpenMP thread and calls the reference routines for
! all arithmetic, with a cell-list radius search in place of coretran (3 x 3
! cells of side >= max r + rtree around each floe, exact distances):
!
!   candidates of i = { j : sqrt((x(i)-x(j))**2 + (y(i)-y(j))**2) <= r(i)+rtree }
!   sorted by ascending distance, ties by ascending index; the first entry
!   (the query floe itself) is dropped, then pairs with i >= j are skipped
!   (stepper.f90:77-87).
!
! Two workload extensions are honoured: a step-1 bond mask (bmask) and
! optional per-floe winds/currents that replace winds_currents' uniform
! values (forcing.f90:24-27).  With bmask all ones and no per-floe forcing the
! run is the reference's own algorithm.
!
! Every entry point is bind(C) so the oracle tests can drive it with ctypes.
! ---------------------------------------------------------------------------

module ref_harness_state
    use iso_c_binding
    implicit none
    integer, allocatable, save :: bmask(:,:)          ! (j,i) 1 = may bond
    integer, save :: use_floe_forcing = 0
    double precision, allocatable, save :: f_ua(:), f_va(:), f_uw(:), f_vw(:)
    ! per-step logs
    integer, save :: nlog = 0, maxlog = 0
    integer, allocatable, save :: log_i(:), log_j(:), log_c(:)
    integer, save :: n_ridge_pair = 0, n_ridge_wall = 0, max_chain = 0
    integer, save :: n_dirty_events = 0
    integer, allocatable, save :: chain(:)
end module ref_harness_state


integer(c_int) function ref_n() bind(C, name="ref_n")
    use iso_c_binding
    implicit none
    include "parameter.h"
    ref_n = n
end function ref_n


! Parameter block, in this order (doubles):
!  1 dt  2 rtree  3 ntree  4 rhoair 5 rhoice 6 rhowater 7 Cdair 8 Cdwater
!  9 Csair 10 Cswater 11 z0w 12 lat 13 e_modul 14 poiss_ratio
! 15 friction_coeff 16 rest_coeff 17 sigmanc_crit 18 eb 19 lambda_rb
! 20 lambda_lb 21 sigmatb_crit 22 sigmacb_crit 23 tau_crit 24 bond_lim
! 25 dmax 26 dtd 27 dth 28 ua 29 va 30 uw 31 vw
! options (ints): 1 slipping 2 cohesion 3 ridging 4 shelter
subroutine ref_setup(prm, opt) bind(C, name="ref_setup")
    use iso_c_binding
    use ref_harness_state
    implicit none
    include "parameter.h"
    include "CB_variables.h"
    include "CB_const.h"
    include "CB_bond.h"
    include "CB_forcings.h"
    include "CB_options.h"
    real(c_double), intent(in) :: prm(31)
    integer(c_int), intent(in) :: opt(4)

    call get_default

    dt = prm(1); rtree = prm(2); ntree = prm(3)
    rhoair = prm(4); rhoice = prm(5); rhowater = prm(6)
    Cdair = prm(7); Cdwater = prm(8); Csair = prm(9); Cswater = prm(10)
    z0w = prm(11); lat = prm(12)
    e_modul = prm(13); poiss_ratio = prm(14); friction_coeff = prm(15)
    rest_coeff = prm(16); sigmanc_crit = prm(17)
    eb = prm(18); lambda_rb = prm(19); lambda_lb = prm(20)
    sigmatb_crit = prm(21); sigmacb_crit = prm(22); tau_crit = prm(23)
    bond_lim = prm(24); dmax = prm(25); dtd = prm(26); dth = prm(27)
    ua = prm(28); va = prm(29); uw = prm(30); vw = prm(31)
    slipping = opt(1) /= 0
    cohesion = opt(2) /= 0
    ridging  = opt(3) /= 0
    shelter  = opt(4) /= 0

    ! derived constants, as read_namelist recomputes them (par_get.f90:263-266)
    ec = e_modul / ( 2 * ( 1 - poiss_ratio ** 2 ) )
    gc = e_modul / ( 4 * ( 1 - poiss_ratio ) * ( 2 + poiss_ratio ) )
    beta =  1d1 * log(rest_coeff) / sqrt( log(rest_coeff) ** 2 + pi ** 2 )

    if (allocated(bmask)) deallocate(bmask)
    allocate(bmask(n, n))
    bmask = 1
    if (allocated(chain)) deallocate(chain)
    allocate(chain(n))
    use_floe_forcing = 0
end subroutine ref_setup


! Reference defaults straight from get_default (par_get.f90:1-111).
subroutine ref_defaults(prm, opt) bind(C, name="ref_defaults")
    use iso_c_binding
    implicit none
    include "parameter.h"
    include "CB_variables.h"
    include "CB_const.h"
    include "CB_bond.h"
    include "CB_forcings.h"
    include "CB_options.h"
    real(c_double), intent(out) :: prm(34)
    integer(c_int), intent(out) :: opt(4)
    call get_default
    prm(1) = dt; prm(2) = rtree; prm(3) = ntree
    prm(4) = rhoair; prm(5) = rhoice; prm(6) = rhowater
    prm(7) = Cdair; prm(8) = Cdwater; prm(9) = Csair; prm(10) = Cswater
    prm(11) = z0w; prm(12) = lat
    prm(13) = e_modul; prm(14) = poiss_ratio; prm(15) = friction_coeff
    prm(16) = rest_coeff; prm(17) = sigmanc_crit
    prm(18) = eb; prm(19) = lambda_rb; prm(20) = lambda_lb
    prm(21) = sigmatb_crit; prm(22) = sigmacb_crit; prm(23) = tau_crit
    prm(24) = bond_lim; prm(25) = dmax; prm(26) = dtd; prm(27) = dth
    prm(28) = ua; prm(29) = va; prm(30) = uw; prm(31) = vw
    prm(32) = ec; prm(33) = gc; prm(34) = beta
    opt(1) = merge(1, 0, slipping); opt(2) = merge(1, 0, cohesion)
    opt(3) = merge(1, 0, ridging);  opt(4) = merge(1, 0, shelter)
end subroutine ref_defaults


subroutine ref_derived(out) bind(C, name="ref_derived")
    use iso_c_binding
    implicit none
    include "parameter.h"
    include "CB_const.h"
    real(c_double), intent(out) :: out(4)
    out(1) = ec; out(2) = gc; out(3) = beta; out(4) = pi
end subroutine ref_derived


! Initial state as ini_get.f90:145-247 leaves it (velocities given).
subroutine ref_set_state(xx, yy, rr, hh, th, om, uu, vv) bind(C, name="ref_set_state")
    use iso_c_binding
    implicit none
    include "parameter.h"
    include "CB_variables.h"
    include "CB_const.h"
    include "CB_bond.h"
    include "CB_forcings.h"
    include "CB_diagnostics.h"
    include "CB_mpi.h"
    real(c_double), intent(in) :: xx(n), yy(n), rr(n), hh(n), th(n), om(n), uu(n), vv(n)
    integer :: i

    x = xx; y = yy; r = rr; h = hh; theta = th; omega = om; u = uu; v = vv
    tfx = 0d0; tfy = 0d0; fcx = 0d0; fcy = 0d0; fbx = 0d0; fby = 0d0
    m = 0d0; mc = 0d0; mb = 0d0
    sigxx = 0d0; sigyy = 0d0; sigxy = 0d0; sigyx = 0d0
    tp = 0d0; tac = 0d0; tab = 0d0; pc = 0d0; pb = 0d0
    fcn = 0d0; fct = 0d0; fcr = 0d0; mbb = 0d0; mcc = 0d0; fbn = 0d0; fbt = 0d0
    thetarelc = 0d0; thetarelb = 0d0; deltat = 0d0; deltanb = 0d0; deltatb = 0d0
    hsfa = 1d0; hsfw = 1d0; ac = 0d0
    theta_bc1 = 0d0; theta_bc2 = 0d0; deltat_bc1 = 0d0; deltat_bc2 = 0d0
    ua_i = 0d0; va_i = 0d0; uw_i = 0d0; vw_i = 0d0
    bond = 0; damageb = 0d0
    rb = 0d0; hb = 0d0; lb = 0d0; sb = 0d0; ib = 0d0
    local_hsfa_min = 0d0; local_hsfw_min = 0d0
    hsfa_min_r = 0d0; hsfw_min_r = 0d0
    do i = 1, n
        call floe_properties(i)
    end do
end subroutine ref_set_state


subroutine ref_set_bondmask(mask) bind(C, name="ref_set_bondmask")
    use iso_c_binding
    use ref_harness_state
    implicit none
    include "parameter.h"
    integer(c_int8_t), intent(in) :: mask(n, n)   ! mask(j,i), 1-based
    bmask = int(mask)
end subroutine ref_set_bondmask


subroutine ref_set_floe_forcing(a, b, c, d) bind(C, name="ref_set_floe_forcing")
    use iso_c_binding
    use ref_harness_state
    implicit none
    include "parameter.h"
    real(c_double), intent(in) :: a(n), b(n), c(n), d(n)
    if (.not. allocated(f_ua)) then
        allocate(f_ua(n), f_va(n), f_uw(n), f_vw(n))
    end if
    f_ua = a; f_va = b; f_uw = c; f_vw = d
    use_floe_forcing = 1
end subroutine ref_set_floe_forcing


subroutine ref_set_log_capacity(cap) bind(C, name="ref_set_log_capacity")
    use iso_c_binding
    use ref_harness_state
    implicit none
    integer(c_int), value :: cap
    if (allocated(log_i)) deallocate(log_i, log_j, log_c)
    maxlog = cap
    allocate(log_i(max(cap, 1)), log_j(max(cap, 1)), log_c(max(cap, 1)))
    nlog = 0
end subroutine ref_set_log_capacity


! One step with stepper.f90's ordering for one rank and one thread.
subroutine ref_step(tstep, restart) bind(C, name="ref_step")
    use iso_c_binding
    use ref_harness_state
    implicit none
    include "parameter.h"
    include "CB_variables.h"
    include "CB_const.h"
    include "CB_bond.h"
    include "CB_forcings.h"
    include "CB_options.h"
    include "CB_mpi.h"
    include "CB_diagnostics.h"
    integer(c_int), value :: tstep, restart

    integer :: i, j, k, c, a, cnt
    double precision :: rad, d, td
    integer, allocatable :: cj(:)
    double precision, allocatable :: cd(:), thr_a(:), thr_w(:)
    double precision :: hi0, hj0
    integer :: tj, ci, cjj
    integer :: gx, gy, cx, cy, bx, by
    double precision :: cs
    integer, allocatable :: head(:), nxt(:)

    allocate(cj(n), cd(n), thr_a(n), thr_w(n))
    ! cell list of the step's start positions: side >= max r + rtree, so the
    ! 3 x 3 cells around a floe hold every floe within its query radius
    ! (cells are clamped into the grid; distances are exact)
    cs = maxval(r) + rtree
    gx = max(1, int(nx / cs)); gy = max(1, int(ny / cs))
    do while (dble(gx) * dble(gy) > 4d0 * n + 64d0)
        gx = max(1, (gx + 1) / 2); gy = max(1, (gy + 1) / 2)
    end do
    allocate(head(gx * gy), nxt(n))
    head = 0
    do j = n, 1, -1
        cx = min(max(int(floor(x(j) / (nx / gx))), 0), gx - 1)
        cy = min(max(int(floor(y(j) / (ny / gy))), 0), gy - 1)
        nxt(j) = head(cy * gx + cx + 1)
        head(cy * gx + cx + 1) = j
    end do
    thr_a = 1d0
    thr_w = 1d0
    nlog = 0
    n_ridge_pair = 0; n_ridge_wall = 0; max_chain = 0; n_dirty_events = 0
    chain = 0

    call reset_forces
    if ( shelter .eqv. .true. ) then
        call reset_shelter
    end if

    do i = 1, n
        call winds_currents(i)
        if (use_floe_forcing /= 0) then
            ua_i(i) = f_ua(i); va_i(i) = f_va(i)
            uw_i(i) = f_uw(i); vw_i(i) = f_vw(i)
        end if

        ! radius search around the current position and radius of i
        rad = r(i) + rtree
        cnt = 0
        cx = min(max(int(floor(x(i) / (nx / gx))), 0), gx - 1)
        cy = min(max(int(floor(y(i) / (ny / gy))), 0), gy - 1)
        do by = max(cy - 1, 0), min(cy + 1, gy - 1)
            do bx = max(cx - 1, 0), min(cx + 1, gx - 1)
                j = head(by * gx + bx + 1)
                do while (j /= 0)
                    d = sqrt( ( x(i) - x(j) ) ** 2 + ( y(i) - y(j) ) ** 2 )
                    if (d <= rad) then
                        cnt = cnt + 1
                        cj(cnt) = j
                        cd(cnt) = d
                    end if
                    j = nxt(j)
                end do
            end do
        end do
        ! insertion sort by (distance, index)
        do a = 2, cnt
            td = cd(a); tj = cj(a)
            c = a - 1
            do while (c >= 1)
                if (cd(c) < td .or. (cd(c) == td .and. cj(c) < tj)) exit
                cd(c + 1) = cd(c); cj(c + 1) = cj(c)
                c = c - 1
            end do
            cd(c + 1) = td; cj(c + 1) = tj
        end do

        do k = 1, cnt - 1
            j = cj(k + 1)
            if (i >= j) cycle

            call rel_pos_vel(j, i)

            if ( tstep == 1 .and. restart /= 1 ) then
                if ( cohesion .eqv. .true. ) then
                    if ( deltan(j, i) >= -bond_lim .and. bmask(j, i) /= 0 ) then
                        bond(j, i) = 1
                        damageb(j, i) = 0d0
                        call bond_properties(j, i)
                    end if
                end if
            end if

            if (chain(i) > 0 .or. chain(j) > 0) n_dirty_events = n_dirty_events + 1
            if ( deltan(j, i) > 0 ) then
                hi0 = h(i); hj0 = h(j)
                call contact_forces(j, i)
                if (h(i) /= hi0 .or. h(j) /= hj0) then
                    n_ridge_pair = n_ridge_pair + 1
                    ci = max(chain(i), chain(j)) + 1
                    if (h(i) /= hi0) chain(i) = ci
                    if (h(j) /= hj0) chain(j) = ci
                    max_chain = max(max_chain, ci)
                end if

                fcx(i) = fcx(i) - fcn(j, i) * cosa(j, i) + fcr(j, i) * sina(j, i)
                fcy(i) = fcy(i) - fcn(j, i) * sina(j, i) - fcr(j, i) * cosa(j, i)
                mc(i) = mc(i) - r(i) * fct(j, i) - mcc(j, i)
                fcx(j) = fcx(j) + fcn(j, i) * cosa(j, i) - fcr(j, i) * sina(j, i)
                fcy(j) = fcy(j) + fcn(j, i) * sina(j, i) + fcr(j, i) * cosa(j, i)
                mc(j) = mc(j) - r(j) * fct(j, i) + mcc(j, i)

                ac(j, i) = delt_ridge(j, i) * min(h(i), h(j))
                tac(i) = tac(i) + ac(j, i)
                pc(i) = pc(i) - fcn(j, i) * ac(j, i)
                tac(j) = tac(j) + ac(j, i)
                pc(j) = pc(j) - fcn(j, i) * ac(j, i)
            else
                call reset_contact(j, i)
            end if

            if (nlog < maxlog) then
                nlog = nlog + 1
                log_i(nlog) = i; log_j(nlog) = j
                log_c(nlog) = 0
                if (deltan(j, i) > 0) log_c(nlog) = 1
            end if

            if ( bond(j, i) == 1 ) then
                call bond_forces_euler(j, i)
                call bond_breaking(j, i)
                if ( bond(j, i) == 1 ) then
                    fbx(i) = fbx(i) + fbn(j, i) * cosa(j, i) - fbt(j, i) * sina(j, i)
                    fby(i) = fby(i) + fbn(j, i) * sina(j, i) + fbt(j, i) * cosa(j, i)
                    mb(i) = mb(i) - lambda_lb * r(i) * fbt(j, i) - mbb(j, i)
                    fbx(j) = fbx(j) - fbn(j, i) * cosa(j, i) + fbt(j, i) * sina(j, i)
                    fby(j) = fby(j) - fbn(j, i) * sina(j, i) - fbt(j, i) * cosa(j, i)
                    mb(j) = mb(j) - lambda_lb * r(j) * fbt(j, i) - mbb(i, j)
                    tab(i) = tab(i) + sb(j, i)
                    pb(i) = pb(i) - fbn(j, i) * sb(j, i)
                    tab(j) = tab(j) + sb(j, i)
                    pb(j) = pb(j) - fbn(j, i) * sb(j, i)
                end if
            else
                call reset_bond(j, i)
            end if

            if ( shelter .eqv. .true. ) then
                call sheltering(j, i)
                thr_a(i) = min(thr_a(i), hsfa(j, i))
                thr_a(j) = min(thr_a(j), hsfa(i, j))
                thr_w(i) = min(thr_w(i), hsfw(j, i))
                thr_w(j) = min(thr_w(j), hsfw(i, j))
            end if

            ! stress diagnostics (stepper.f90:238-271)
            sigxx(i) = sigxx(i) - ( sqrt(fcn(j,i) ** 2 + fct(j,i) ** 2) + &
                        sqrt(fbn(j,i) ** 2 + fbt(j,i) ** 2)) * cosa(j,i) * r(i) * cosa(j,i)
            sigyy(i) = sigyy(i) - ( sqrt(fcn(j,i) ** 2 + fct(j,i) ** 2) + &
                        sqrt(fbn(j,i) ** 2 + fbt(j,i) ** 2)) * sina(j,i) * r(i) * sina(j,i)
            sigxy(i) = sigxy(i) - ( sqrt(fcn(j,i) ** 2 + fct(j,i) ** 2) + &
                        sqrt(fbn(j,i) ** 2 + fbt(j,i) ** 2)) * sina(j,i) * r(i) * cosa(j,i)
            sigyx(i) = sigyx(i) - ( sqrt(fcn(j,i) ** 2 + fct(j,i) ** 2) + &
                        sqrt(fbn(j,i) ** 2 + fbt(j,i) ** 2)) * cosa(j,i) * r(i) * sina(j,i)
            sigxx(j) = sigxx(j) - ( sqrt(fcn(j,i) ** 2 + fct(j,i) ** 2) + &
                        sqrt(fbn(j,i) ** 2 + fbt(j,i) ** 2)) * cosa(j,i) * r(i) * cosa(j,i)
            sigyy(j) = sigyy(j) - ( sqrt(fcn(j,i) ** 2 + fct(j,i) ** 2) + &
                        sqrt(fbn(j,i) ** 2 + fbt(j,i) ** 2)) * sina(j,i) * r(i) * sina(j,i)
            sigxy(j) = sigxy(j) - ( sqrt(fcn(j,i) ** 2 + fct(j,i) ** 2) + &
                        sqrt(fbn(j,i) ** 2 + fbt(j,i) ** 2)) * sina(j,i) * r(i) * cosa(j,i)
            sigyx(j) = sigyx(j) - ( sqrt(fcn(j,i) ** 2 + fct(j,i) ** 2) + &
                        sqrt(fbn(j,i) ** 2 + fbt(j,i) ** 2)) * cosa(j,i) * r(i) * sina(j,i)
        end do

        hi0 = h(i)
        call verify_bc(i)
        if (h(i) /= hi0) then
            n_ridge_wall = n_ridge_wall + 1
        end if
    end do

    ! one thread, one rank: the reductions and collectives are copies
    do i = 1, n
        local_hsfa_min(i) = min(local_hsfa_min(i), thr_a(i))
        local_hsfw_min(i) = min(local_hsfw_min(i), thr_w(i))
    end do
    hsfa_min_r = local_hsfa_min
    hsfw_min_r = local_hsfw_min

    do i = 1, n
        call forcing(i)
        call coriolis(i)
    end do

    fcx_r = fcx; fcy_r = fcy; mc_r = mc
    fbx_r = fbx; fby_r = fby; mb_r = mb
    fwx_r = fwx; fwy_r = fwy; mw_r = mw
    fax_r = fax; fay_r = fay; ma_r = ma
    fcorx_r = fcorx; fcory_r = fcory
    fx_bc_r = fx_bc; fy_bc_r = fy_bc; m_bc_r = m_bc
    sigxx_r = sigxx; sigyy_r = sigyy; sigxy_r = sigxy; sigyx_r = sigyx
    sigxx_bc_r = sigxx_bc; sigyy_bc_r = sigyy_bc; sigxy_bc_r = sigxy_bc; sigyx_bc_r = sigyx_bc
    sigxx_aw_r = sigxx_aw; sigyy_aw_r = sigyy_aw; sigxy_aw_r = sigxy_aw; sigyx_aw_r = sigyx_aw
    tac_r = tac; tab_r = tab; pc_r = pc; pb_r = pb
    ta_bc_r = ta_bc; p_bc_r = p_bc

    ! totals (stepper.f90:312-332)
    do i = 1, n
        tfx_r(i) = fcx_r(i) + fbx_r(i) + fax_r(i) + fwx_r(i) + fcorx_r(i) + fx_bc_r(i)
        tfy_r(i) = fcy_r(i) + fby_r(i) + fay_r(i) + fwy_r(i) + fcory_r(i) + fy_bc_r(i)
        m_r(i) = mc_r(i) + mb_r(i) + ma_r(i) + mw_r(i) + m_bc_r(i)
        tsigxx_r(i) = sigxx_r(i) + sigxx_bc_r(i) + sigxx_aw_r(i)
        tsigyy_r(i) = sigyy_r(i) + sigyy_bc_r(i) + sigyy_aw_r(i)
        tsigxy_r(i) = sigxy_r(i) + sigxy_bc_r(i) + sigxy_aw_r(i)
        tsigyx_r(i) = sigyx_r(i) + sigyx_bc_r(i) + sigyx_aw_r(i)
        tp_r(i) = merge(pc_r(i) / tac_r(i), 0d0, tac_r(i) /= 0d0) &
                + merge(pb_r(i) / tab_r(i), 0d0, tab_r(i) /= 0d0) &
                + merge(p_bc_r(i) / ta_bc_r(i), 0d0, ta_bc_r(i) /= 0d0)
    end do
    tfx = tfx_r; tfy = tfy_r; m = m_r
    tsigxx = tsigxx_r; tsigyy = tsigyy_r; tsigxy = tsigxy_r; tsigyx = tsigyx_r
    tp = tp_r

    call velocity
    call position

    deallocate(cj, cd, thr_a, thr_w, head, nxt)
end subroutine ref_step


subroutine ref_get_state(xx, yy, rr, hh, th, om, uu, vv, ms, ine, ha, hw) &
        bind(C, name="ref_get_state")
    use iso_c_binding
    implicit none
    include "parameter.h"
    include "CB_variables.h"
    real(c_double), intent(out) :: xx(n), yy(n), rr(n), hh(n), th(n), om(n), &
                                   uu(n), vv(n), ms(n), ine(n), ha(n), hw(n)
    xx = x; yy = y; rr = r; hh = h; th = theta; om = omega; uu = u; vv = v
    ms = mass; ine = inertia; ha = hfa; hw = hfw
end subroutine ref_get_state


! Per-floe force record of the last step, columns:
!  1 tfx 2 tfy 3 m 4 fcx 5 fcy 6 mc 7 fbx 8 fby 9 mb 10 fax 11 fay 12 ma
! 13 fwx 14 fwy 15 mw 16 fcorx 17 fcory 18 fx_bc 19 fy_bc 20 m_bc
! 21 hsfa_min 22 hsfw_min 23 tp 24 tsigxx 25 tsigyy 26 tsigxy 27 tsigyx
! 28 tac 29 pc 30 tab 31 pb 32 ta_bc 33 p_bc
subroutine ref_get_forces(f) bind(C, name="ref_get_forces")
    use iso_c_binding
    implicit none
    include "parameter.h"
    include "CB_variables.h"
    include "CB_bond.h"
    include "CB_forcings.h"
    include "CB_mpi.h"
    include "CB_diagnostics.h"
    real(c_double), intent(out) :: f(n, 33)
    f(:,1) = tfx; f(:,2) = tfy; f(:,3) = m
    f(:,4) = fcx; f(:,5) = fcy; f(:,6) = mc
    f(:,7) = fbx; f(:,8) = fby; f(:,9) = mb
    f(:,10) = fax; f(:,11) = fay; f(:,12) = ma
    f(:,13) = fwx; f(:,14) = fwy; f(:,15) = mw
    f(:,16) = fcorx; f(:,17) = fcory
    f(:,18) = fx_bc; f(:,19) = fy_bc; f(:,20) = m_bc
    f(:,21) = hsfa_min_r; f(:,22) = hsfw_min_r; f(:,23) = tp
    f(:,24) = tsigxx; f(:,25) = tsigyy; f(:,26) = tsigxy; f(:,27) = tsigyx
    f(:,28) = tac; f(:,29) = pc; f(:,30) = tab; f(:,31) = pb
    f(:,32) = ta_bc; f(:,33) = p_bc
end subroutine ref_get_forces


! Wall histories: 1 deltat_bc1 2 deltat_bc2 3 theta_bc1 4 theta_bc2
subroutine ref_get_walls(w) bind(C, name="ref_get_walls")
    use iso_c_binding
    implicit none
    include "parameter.h"
    include "CB_variables.h"
    real(c_double), intent(out) :: w(n, 4)
    w(:,1) = deltat_bc1; w(:,2) = deltat_bc2; w(:,3) = theta_bc1; w(:,4) = theta_bc2
end subroutine ref_get_walls


integer(c_int) function ref_get_log(cap, li, lj, lc) bind(C, name="ref_get_log")
    use iso_c_binding
    use ref_harness_state
    implicit none
    integer(c_int), value :: cap
    integer(c_int), intent(out) :: li(cap), lj(cap), lc(cap)
    integer :: k
    do k = 1, min(cap, nlog)
        li(k) = log_i(k); lj(k) = log_j(k); lc(k) = log_c(k)
    end do
    ref_get_log = nlog
end function ref_get_log


subroutine ref_get_stats(s) bind(C, name="ref_get_stats")
    use iso_c_binding
    use ref_harness_state
    implicit none
    integer(c_int), intent(out) :: s(4)
    s(1) = n_ridge_pair; s(2) = n_ridge_wall; s(3) = max_chain; s(4) = n_dirty_events
end subroutine ref_get_stats


! Dense pair-state export for the listed pairs (1-based i<j), columns:
!  1 deltat 2 thetarelc 3 bond 4 damageb 5 deltanb 6 deltatb
!  7 thetarelb(j,i) 8 thetarelb(i,j) 9 rb 10 hb 11 lb 12 sb 13 ib
! 14 fcn 15 fct 16 fbn 17 fbt 18 mcc
subroutine ref_get_pairs(np, pi_, pj_, out) bind(C, name="ref_get_pairs")
    use iso_c_binding
    implicit none
    include "parameter.h"
    include "CB_variables.h"
    include "CB_bond.h"
    integer(c_int), value :: np
    integer(c_int), intent(in) :: pi_(np), pj_(np)
    real(c_double), intent(out) :: out(np, 18)
    integer :: k, i, j
    do k = 1, np
        i = pi_(k); j = pj_(k)
        out(k,1) = deltat(j,i); out(k,2) = thetarelc(j,i)
        out(k,3) = dble(bond(j,i)); out(k,4) = damageb(j,i)
        out(k,5) = deltanb(j,i); out(k,6) = deltatb(j,i)
        out(k,7) = thetarelb(j,i); out(k,8) = thetarelb(i,j)
        out(k,9) = rb(j,i); out(k,10) = hb(j,i); out(k,11) = lb(j,i)
        out(k,12) = sb(j,i); out(k,13) = ib(j,i)
        out(k,14) = fcn(j,i); out(k,15) = fct(j,i)
        out(k,16) = fbn(j,i); out(k,17) = fbt(j,i); out(k,18) = mcc(j,i)
    end do
end subroutine ref_get_pairs


! Bonded pair list (1-based, i<j): returns the count, fills up to cap.
integer(c_int) function ref_get_bonds(cap, bi, bj) bind(C, name="ref_get_bonds")
    use iso_c_binding
    implicit none
    include "parameter.h"
    include "CB_bond.h"
    integer(c_int), value :: cap
    integer(c_int), intent(out) :: bi(cap), bj(cap)
    integer :: i, j, c
    c = 0
    do i = 1, n
        do j = i + 1, n
            if (bond(j, i) == 1) then
                c = c + 1
                if (c <= cap) then
                    bi(c) = i; bj(c) = j
                end if
            end if
        end do
    end do
    ref_get_bonds = c
end function ref_get_bonds
