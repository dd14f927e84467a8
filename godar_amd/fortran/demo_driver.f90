!=======================================================================
! Stand-in for the reference's time loop (model/src/godar.f90:191-213)
! that drives the GPU library through stepper_gpu, the drop-in for
! stepper.  It is compiled against the reference's COMMON blocks (with a
! parameter.h for its n, nx, ny), like stepper_gpu itself.
!
!     demo_driver <input> <output> [expno]
!
! With expno, the reference's own output routine sea_ice_post
! (diagnostics_outputs.f90, compiled from the reference sources) writes
! output/<name>.<expno> on the output steps, as godar.f90:195-197 does.
!
! input (stream, native byte order): int32 nt, restart, comp, precision;
! 33 doubles dt rtree rhoair rhoice rhowater Cdair Cdwater Csair Cswater
! z0w lat e_modul poiss_ratio friction_coeff rest_coeff sigmanc_crit eb
! lambda_rb lambda_lb sigmatb_crit sigmacb_crit tau_crit bond_lim dmax dtd
! dth ua va uw vw ntree (as double) unused unused; int32 slipping cohesion
! ridging shelter; then x y u v theta omega r h (n doubles each).
! output: x y u v theta omega r h mass inertia tfx tfy m (n doubles each),
! int32 number of bonds, then (i, j) 1-based pairs with bond(j,i) == 1.
!=======================================================================
program demo_driver
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

    character(len=512) :: fin, fout
    character(len=4) :: expno
    logical :: post
    integer(c_int32_t) :: hdr(4), opt(4), nbonds
    real(c_double) :: pv(33)
    integer :: tstep, restart_, ii, jj, u_in, u_out

    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    post = command_argument_count() >= 3
    if (post) call get_command_argument(3, expno)
    open(newunit=u_in, file=trim(fin), access='stream', form='unformatted', status='old')
    read(u_in) hdr, pv, opt
    read(u_in) x, y, u, v, theta, omega, r, h
    close(u_in)

    nt = dble(hdr(1)); restart_ = hdr(2); comp = dble(hdr(3))
    godar_stepper_precision = hdr(4)
    dt = pv(1); rtree = pv(2); rhoair = pv(3); rhoice = pv(4); rhowater = pv(5)
    Cdair = pv(6); Cdwater = pv(7); Csair = pv(8); Cswater = pv(9); z0w = pv(10); lat = pv(11)
    e_modul = pv(12); poiss_ratio = pv(13); friction_coeff = pv(14); rest_coeff = pv(15)
    sigmanc_crit = pv(16); eb = pv(17); lambda_rb = pv(18); lambda_lb = pv(19)
    sigmatb_crit = pv(20); sigmacb_crit = pv(21); tau_crit = pv(22); bond_lim = pv(23)
    dmax = pv(24); dtd = pv(25); dth = pv(26); ua = pv(27); va = pv(28); uw = pv(29); vw = pv(30)
    ntree = pv(31)
    pi = 4d0 * datan(1d0)        ! par_get.f90:35 (the output routine reads it)
    slipping = opt(1) /= 0; cohesion = opt(2) /= 0; ridging = opt(3) /= 0; shelter = opt(4) /= 0
    bond = 0
    damageb = 0d0

    do tstep = 1, int(nt)
        call stepper_gpu(tstep, restart_)
        if (post .and. (modulo(tstep, int(comp)) .eq. 0 .or. tstep + restart_ == 1)) then
            call sea_ice_post(tstep, expno)
        end if
    end do

    nbonds = 0
    do ii = 1, n
        do jj = ii + 1, n
            if (bond(jj, ii) == 1) nbonds = nbonds + 1
        end do
    end do
    open(newunit=u_out, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(u_out) x, y, u, v, theta, omega, r, h, mass, inertia, tfx, tfy, m
    write(u_out) nbonds
    do ii = 1, n
        do jj = ii + 1, n
            if (bond(jj, ii) == 1) write(u_out) int(ii, c_int32_t), int(jj, c_int32_t)
        end do
    end do
    close(u_out)
end program demo_driver
