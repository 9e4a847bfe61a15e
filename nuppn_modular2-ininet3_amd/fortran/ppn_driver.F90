! ppn_driver - a one-zone ppn cycle loop that calls the MI355X library through the nuppn_hip
! module, the way frames/ppn/source/ppn.F90:289-313 calls evaluate_all_rates + integrate_network
! and dispatches on ierr, with mppnp's one-call-per-cycle batch form as a second mode.
!
! Usage:  ppn_driver zone  net.bin traj.bin out.bin
!         ppn_driver batch net.bin zones.bin out.bin
!
! net.bin (stream, native endianness): nsp, nre, igamma (int32); k(nre, 8), rtype(nre) (int32);
!   coef(nre, 7), aw(nsp), zw(nsp) (float64).
! zone mode: traj.bin = ncyc (int32), x0(nsp), then per cycle t9_0, t9_1, rho_0, rho_1, dtwant
!   (float64).  out.bin = per cycle rc, ierr, nsubt, titers, n_newton (int32, n_newton as
!   int64), then the final yps(nsp) and ye (float64).
! batch mode: zones.bin = nz (int32), yps(nsp, nz), t9_0(nz), t9_1(nz), rho_0(nz), rho_1(nz),
!   dtwant(nz) (float64).  out.bin = rc, ierr(nz), nsubt(nz), n_newton(nz) (int32, n_newton as
!   int64), yps(nsp, nz), ye(nz).
program ppn_driver
  use iso_c_binding
  use nuppn_hip
  implicit none
  character(len=512) :: mode, fnet, fin, fout
  integer(c_int) :: nsp, nre, igamma, status, ncyc, icyc, ierr, rc, nz, i
  integer(c_int), allocatable, target :: k(:, :), rtype(:), ierrz(:)
  real(c_double), allocatable, target :: coef(:, :), aw(:), zw(:), yps(:), ypsz(:, :)
  real(c_double), allocatable :: t9_0(:), t9_1(:), rho_0(:), rho_1(:), dtw(:), yez(:)
  type(nuppn_zone_stats), allocatable, target :: stz(:)
  real(c_double) :: t9a, t9b, rhoa, rhob, dtwant, ye
  type(nuppn_network) :: net
  type(nuppn_knobs) :: knobs
  type(nuppn_zone_stats) :: st
  type(c_ptr) :: ctx
  integer :: u

  call get_command_argument(1, mode)
  call get_command_argument(2, fnet)
  call get_command_argument(3, fin)
  call get_command_argument(4, fout)

  open(newunit=u, file=trim(fnet), access='stream', form='unformatted', status='old')
  read(u) nsp, nre, igamma
  allocate(k(nre, 8), rtype(nre), coef(nre, 7), aw(nsp), zw(nsp))
  read(u) k, rtype, coef, aw, zw
  close(u)

  ! solver_init + makemap: the context (knobs as readsolverinput's defaults, static network)
  net%nsp = nsp
  net%nre = nre
  net%k = c_loc(k)
  net%rtype = c_loc(rtype)
  net%coef = c_loc(coef)
  net%aw = c_loc(aw)
  net%zw = c_loc(zw)
  net%igamma = igamma
  call nuppn_default_knobs(knobs)
  ctx = nuppn_ctx_create(net, knobs, status)
  if (.not. c_associated(ctx)) then
     write(*, '(a, i0)') 'ppn_driver: nuppn_ctx_create failed, status ', status
     stop 2
  end if

  select case (trim(mode))
  case ('zone')
     allocate(yps(nsp))
     open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
     read(u) ncyc
     read(u) yps
     open(unit=21, file=trim(fout), access='stream', form='unformatted', status='replace')
     ye = 0.5d0
     do icyc = 1, ncyc
        read(u) t9a, t9b, rhoa, rhob, dtwant
        ! integrate_network's body (solver/source/solver.F90:53-238) becomes this call
        rc = nuppn_integrate_zone(ctx, yps, t9a, t9b, rhoa, rhob, ye, dtwant, ierr, st)
        if (rc /= NUPPN_OK) ierr = NUPPN_BAD_STEP
        write(21) rc, ierr, st%nsubt, st%titers, st%n_newton
        ! ppn.F90:303-313 dispatches on ierr; this driver stops at the first failure
        select case (ierr)
        case (NUPPN_OK)
        case default
           write(*, '(a, i0, a, i0)') 'ppn_driver: cycle ', icyc, ' ierr ', ierr
           exit
        end select
     end do
     close(u)
     write(21) yps, ye
     close(21)
     write(*, '(a, i0, a)') 'ppn_driver: ', ncyc, ' cycles done'
  case ('batch')
     open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
     read(u) nz
     allocate(ypsz(nsp, nz), t9_0(nz), t9_1(nz), rho_0(nz), rho_1(nz), dtw(nz), yez(nz), &
              ierrz(nz), stz(nz))
     read(u) ypsz, t9_0, t9_1, rho_0, rho_1, dtw
     close(u)
     yez = 0.5d0
     ! mppnp's worker loop (frames/mppnp/source/mppnp.F90:713-753) as one call over the cycle
     rc = nuppn_integrate_batch_mixed(ctx, nz, ypsz, t9_0, t9_1, rho_0, rho_1, yez, dtw, &
                                      ierrz, c_loc(stz), c_null_ptr)
     open(unit=21, file=trim(fout), access='stream', form='unformatted', status='replace')
     write(21) rc, ierrz, (stz(i)%nsubt, i = 1, nz), (stz(i)%n_newton, i = 1, nz)
     write(21) ypsz, yez
     close(21)
     write(*, '(a, i0, a, i0)') 'ppn_driver: batch of ', nz, ' zones, rc ', rc
  case default
     write(*, '(a)') 'usage: ppn_driver zone|batch net.bin in.bin out.bin'
     stop 1
  end select
  call nuppn_ctx_destroy(ctx)
end program ppn_driver
