!=======================================================================
! Drop-in replacement of the reference's step subroutine
!     subroutine stepper (tstep, restart)               model/src/stepper.f90:1
! called from the time loop at model/src/godar.f90:193.  Same arguments, same
! COMMON blocks (it includes model/inc/*.h, so it is compiled inside the
! reference tree, next to stepper.f90):
!
!     call stepper_gpu (tstep, restart)
!
! The floe state lives on the GPU between calls.  The first call creates a
! context from the COMMON parameters and the COMMON floe state (and, when
! restart == 1, the COMMON bond matrices).  Every call advances one step.  The
! COMMON floe state, forces and bond matrices are refreshed on the steps where
! the driver writes output, modulo(tstep, comp) == 0 .or. tstep + restart == 1
! (godar.f90:195), the only steps on which it reads them; set
! godar_sync_every = .true. to refresh after every step instead.
!
! What is refreshed: x y u v theta omega r h mass inertia hfa hfw; tfx tfy m
! fcx fcy mc fbx fby mb fax fay ma fwx fwy mw fcorx fcory fx_bc fy_bc m_bc tp
! tsigxx tsigyy tsigxy tsigyx; bond(j,i) and damageb(j,i) of the bonded pairs
! (others 0); deltan(j,i) > 0 marks the contact pairs of the step (value
! r(i)+r(j)-dist at the refreshed positions, at least tiny(1d0)), which is all
! the output routine reads of it (diagnostics_outputs.f90:164).  Forces and
! stresses are those of the last step (the library records them every step).
!=======================================================================
subroutine stepper_gpu (tstep, restart)

    use iso_c_binding
    use godar_gpu

    implicit none

    include "parameter.h"
    include "CB_variables.h"
    include "CB_const.h"
    include "CB_bond.h"
    include "CB_forcings.h"
    include "CB_options.h"
    include "CB_diagnostics.h"

    integer, intent(in) :: tstep, restart

    type(c_ptr), save :: ctx = c_null_ptr
    integer(c_int32_t), save :: mem = -1
    type(godar_params) :: prm
    integer(c_int) :: rc
    integer :: k, q, nbd, ii, jj
    integer(c_int32_t) :: cnt
    real(c_double), allocatable :: fo(:, :), st(:, :)
    integer(c_int32_t), allocatable :: pi_(:), pj_(:)
    integer(c_int32_t) :: none_(0)
    integer, allocatable :: bi_(:), bj_(:)
    double precision :: dd

    if (.not. c_associated(ctx)) then
        rc = godar_params_default(prm)
        prm%dt = dt; prm%rtree = rtree; prm%nx = nx; prm%ny = ny
        prm%rhoair = rhoair; prm%rhoice = rhoice; prm%rhowater = rhowater
        prm%Cdair = Cdair; prm%Cdwater = Cdwater; prm%Csair = Csair; prm%Cswater = Cswater
        prm%z0w = z0w; prm%lat = lat
        prm%e_modul = e_modul; prm%poiss_ratio = poiss_ratio
        prm%friction_coeff = friction_coeff; prm%rest_coeff = rest_coeff
        prm%sigmanc_crit = sigmanc_crit
        prm%eb = eb; prm%lambda_rb = lambda_rb; prm%lambda_lb = lambda_lb
        prm%sigmatb_crit = sigmatb_crit; prm%sigmacb_crit = sigmacb_crit; prm%tau_crit = tau_crit
        prm%bond_lim = bond_lim; prm%dmax = dmax; prm%dtd = dtd; prm%dth = dth
        prm%ua = ua; prm%va = va; prm%uw = uw; prm%vw = vw
        prm%bond_frac = 1d0
        prm%ntree = int(ntree, c_int32_t)
        prm%slipping = merge(1, 0, slipping); prm%cohesion = merge(1, 0, cohesion)
        prm%ridging = merge(1, 0, ridging); prm%shelter = merge(1, 0, shelter)
        prm%restart = int(restart, c_int32_t)
        rc = godar_create(int(n, c_int64_t), 1_c_int32_t, godar_stepper_precision, ctx)
        if (rc /= GODAR_OK) then
            write(*, '(A, I0)') 'stepper_gpu: godar_create failed, status ', rc
            stop 1
        end if
        call godar_check(godar_record_forces(ctx, 1_c_int32_t), ctx, 'record_forces')
        call godar_check(godar_member_add(ctx, prm, int(n, c_int32_t), x, y, u, v, theta, omega, &
                                          r, h, int(tstep, c_int32_t), mem), ctx, 'member_add')
        if (restart == 1) then
            ! bonds and their state from the COMMON matrices (column i holds
            ! the pairs (j, i), j > i)
            nbd = 0
            do ii = 1, n
                do jj = ii + 1, n
                    if (bond(jj, ii) == 1 .or. damageb(jj, ii) /= 0d0) nbd = nbd + 1
                end do
            end do
            allocate(bi_(max(nbd, 1)), bj_(max(nbd, 1)), st(GODAR_NBONDSTATE, max(nbd, 1)))
            q = 0
            do ii = 1, n
                do jj = ii + 1, n
                    if (bond(jj, ii) == 1 .or. damageb(jj, ii) /= 0d0) then
                        q = q + 1
                        bi_(q) = ii
                        bj_(q) = jj
                        st(:, q) = [dble(bond(jj, ii)), damageb(jj, ii), deltanb(jj, ii), &
                                    deltatb(jj, ii), thetarelb(jj, ii), thetarelb(ii, jj), &
                                    rb(jj, ii), hb(jj, ii), lb(jj, ii), sb(jj, ii), ib(jj, ii)]
                    end if
                end do
            end do
            call godar_check(godar_set_bonds(ctx, mem, int(nbd, c_int32_t), bi_, bj_, st), ctx, &
                             'set_bonds')
            deallocate(bi_, bj_, st)
        end if
    end if

    call godar_check(godar_step(ctx, 1_c_int32_t), ctx, 'step')

    if (.not. (godar_sync_every .or. modulo(tstep, int(comp)) == 0 .or. tstep + restart == 1)) return

    ! floe state
    call godar_check(godar_get_floes(ctx, mem, x, y, u, v, theta, omega, r, h, mass, inertia, &
                                     hfa, hfw), ctx, 'get_floes')
    ! forces of the step: the C (33, n) row-major block is fo(n, 33)
    allocate(fo(n, GODAR_NFORCES))
    call godar_check(godar_get_forces(ctx, mem, fo), ctx, 'get_forces')
    tfx = fo(:, GF_TFX); tfy = fo(:, GF_TFY); m = fo(:, GF_M)
    fcx = fo(:, GF_FCX); fcy = fo(:, GF_FCY); mc = fo(:, GF_MC)
    fbx = fo(:, GF_FBX); fby = fo(:, GF_FBY); mb = fo(:, GF_MB)
    fax = fo(:, GF_FAX); fay = fo(:, GF_FAY); ma = fo(:, GF_MA)
    fwx = fo(:, GF_FWX); fwy = fo(:, GF_FWY); mw = fo(:, GF_MW)
    fcorx = fo(:, GF_FCORX); fcory = fo(:, GF_FCORY)
    fx_bc = fo(:, GF_FX_BC); fy_bc = fo(:, GF_FY_BC); m_bc = fo(:, GF_M_BC)
    tp = fo(:, GF_TP)
    tsigxx = fo(:, GF_TSIGXX); tsigyy = fo(:, GF_TSIGYY)
    tsigxy = fo(:, GF_TSIGXY); tsigyx = fo(:, GF_TSIGYX)
    deallocate(fo)

    ! bonded pairs and damage
    bond = 0
    damageb = 0d0
    call godar_check(godar_get_pairs(ctx, mem, GODAR_PAIRS_BOND, none_, none_, cnt), &
                     ctx, 'get_pairs')
    allocate(pi_(max(cnt, 1)), pj_(max(cnt, 1)), st(GODAR_NBONDSTATE, max(cnt, 1)))
    call godar_check(godar_get_pairs(ctx, mem, GODAR_PAIRS_BOND, pi_, pj_, cnt), ctx, 'get_pairs')
    call godar_check(godar_get_bond_state(ctx, mem, cnt, st, cnt), ctx, 'get_bond_state')
    do k = 1, int(cnt)
        bond(pj_(k), pi_(k)) = 1
        damageb(pj_(k), pi_(k)) = st(2, k)
    end do
    deallocate(pi_, pj_, st)

    ! contact pairs of the step
    deltan = 0d0
    call godar_check(godar_get_pairs(ctx, mem, GODAR_PAIRS_CONTACT, none_, none_, cnt), &
                     ctx, 'get_pairs')
    allocate(pi_(max(cnt, 1)), pj_(max(cnt, 1)))
    call godar_check(godar_get_pairs(ctx, mem, GODAR_PAIRS_CONTACT, pi_, pj_, cnt), ctx, 'get_pairs')
    do k = 1, int(cnt)
        ii = pi_(k)
        jj = pj_(k)
        dd = r(ii) + r(jj) - sqrt((x(ii) - x(jj)) ** 2 + (y(ii) - y(jj)) ** 2)
        deltan(jj, ii) = max(dd, tiny(1d0))
    end do
    deallocate(pi_, pj_)

end subroutine stepper_gpu
