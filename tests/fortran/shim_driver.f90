! Test driver for the Fortran drop-in module mod_sod2d_hip: a caller written
! the way the reference's solver calls its routines, through the reference
! names and dummy lists only.
!
!   shim_driver tables <out>      host-only: p=3 tables and a 2^3 periodic
!                                 box mesh, written as raw stream data
!   shim_driver path <in> <out>   GPU: reads a mesh + state, registers the
!                                 host arrays (enter data), runs the init
!                                 routines, one call of each hot-path routine
!                                 and one rk_4_main step, writes the results
program shim_driver
   use iso_c_binding
   use mod_sod2d_hip
   implicit none
   character(len=512) :: mode, fin, fout

   call get_command_argument(1, mode)
   if (trim(mode) == 'tables') then
      call get_command_argument(2, fout)
      call run_tables(trim(fout))
   else if (trim(mode) == 'path') then
      call get_command_argument(2, fin)
      call get_command_argument(3, fout)
      call run_path(trim(fin), trim(fout))
   else
      print '(a)', 'usage: shim_driver tables <out> | path <in> <out>'
      error stop 2
   end if

contains

   subroutine run_tables(fout)
      character(*), intent(in) :: fout
      integer(4) :: atoIJK(64), invAtoIJK(4, 4, 4), gI(64), gJ(64), gK(64)
      real(8) :: xgp(64, 3), wgp(64), Ngp(64, 64), dNgp(3, 64, 64), dlxigp_ip(64, 3, 4)
      integer(4), allocatable :: connecOrig(:, :), connec(:, :), lpoin_w(:)
      real(8), allocatable :: coord(:, :)
      integer :: u
      call sod2d_hex64_tables(atoIJK, xgp, wgp, Ngp, dNgp, dlxigp_ip, invAtoIJK, gI, gJ, gK)
      allocate (connecOrig(8, 64), connec(8, 64), lpoin_w(216), coord(343, 3))
      call sod2d_structured_box_mesh(2, 2, 2, 6.283185307179586d0, 6.283185307179586d0, 6.283185307179586d0, &
                                     0.0d0, connecOrig, connec, lpoin_w, coord)
      open (newunit=u, file=fout, access='stream', form='unformatted', status='replace')
      write (u) atoIJK, xgp, wgp, Ngp, dNgp, dlxigp_ip, invAtoIJK, gI, gJ, gK, connecOrig, connec, lpoin_w, coord
      close (u)
   end subroutine

   ! registers a host array with the library (enter data copyin/create)
   subroutine reg(p, bytes, copyin)
      type(c_ptr), intent(in) :: p
      integer(8), intent(in) :: bytes
      logical, intent(in) :: copyin
      call sod2d_enter_data(p, bytes, copyin)
   end subroutine

   subroutine run_path(fin, fout)
      character(*), intent(in) :: fin, fout
      integer(4) :: nelem, npoin, npoin_w, u
      integer(4), target :: atoIJK(64), invAtoIJK(4, 4, 4), gI(64), gJ(64), gK(64)
      real(8), target :: xgp(64, 3), wgp(64), Ngp(64, 64), dNgp(3, 64, 64), dlxigp_ip(64, 3, 4)
      integer(4), allocatable, target :: connecOrig(:, :), connec(:, :), lpoin_w(:), point2elem(:), lnbn_nodes(:)
      real(8), allocatable, target :: coord(:, :), rho(:, :), uu(:, :, :), q(:, :, :), pr(:, :), E(:, :)
      real(8), allocatable, target :: Tem(:, :), e_int(:, :), eta(:, :), csound(:), machno(:), mu_fluid(:)
      real(8), allocatable, target :: mu_factor(:), mu_e(:, :), mu_sgs(:, :), mue_l(:, :)
      real(8), allocatable, target :: gpvol(:, :, :), He(:, :, :, :), Ml(:), helem(:), helem_l(:, :)
      real(8), allocatable, target :: Rmass(:), Rmom(:, :), Rener(:), Rconv(:), fq(:, :)
      real(8), allocatable, target :: Dmass(:), Dmom(:, :), Dener(:)
      real(8), allocatable, target :: kres(:), etot(:), au(:, :), ax1(:), ax2(:), ax3(:)
      real(8), target :: Rgas, gamma_gas, Cp, Prt, dt, dt_cfl, cfl, EK
      integer(8), parameter :: i4 = 4, r8 = 8
      type(sod2d_numerical_params) :: prm

      open (newunit=u, file=fin, access='stream', form='unformatted', status='old')
      read (u) nelem, npoin, npoin_w
      allocate (connecOrig(nelem, 64), connec(nelem, 64), lpoin_w(npoin_w), coord(npoin, 3))
      allocate (rho(npoin, 3), uu(npoin, 3, 3), q(npoin, 3, 3), pr(npoin, 3), E(npoin, 3), Tem(npoin, 2))
      allocate (e_int(npoin, 2), eta(npoin, 3), csound(npoin), machno(npoin), mu_fluid(npoin), mu_factor(npoin))
      allocate (mu_e(nelem, 64), mu_sgs(nelem, 64), mue_l(nelem, 64))
      rho = 0; uu = 0; q = 0; pr = 0; E = 0; Tem = 0; e_int = 0; eta = 0
      read (u) connecOrig, connec, lpoin_w, coord
      read (u) rho(:, 2), uu(:, :, 2), q(:, :, 2), pr(:, 2), E(:, 2), Tem(:, 2), e_int(:, 2), eta(:, 2)
      read (u) csound, machno, mu_fluid, mu_factor, mu_e
      read (u) Rgas, gamma_gas, Cp, Prt, dt
      close (u)
      mu_sgs = 0; mue_l = 0
      allocate (point2elem(npoin), lnbn_nodes(npoin))
      point2elem = 0; lnbn_nodes = 0
      allocate (gpvol(1, 64, nelem), He(3, 3, 64, nelem), Ml(npoin), helem(nelem), helem_l(nelem, 64))
      allocate (Rmass(npoin), Rmom(npoin, 3), Rener(npoin), Rconv(npoin), fq(npoin, 3))
      allocate (Dmass(npoin), Dmom(npoin, 3), Dener(npoin))
      allocate (kres(npoin), etot(npoin), au(npoin, 3), ax1(npoin), ax2(npoin), ax3(npoin))
      kres = 0; etot = 0; au = 0; ax1 = 0; ax2 = 0; ax3 = 0
      fq(:, 1) = uu(:, 1, 2)*eta(:, 2)
      fq(:, 2) = uu(:, 2, 2)*eta(:, 2)
      fq(:, 3) = uu(:, 3, 2)*eta(:, 2)

      call sod2d_init()
      ! the case sets nscbc_Rgas_inf = Rgas (TGVSolver.f90:104)
      call sod2d_get_numerical_params(prm)
      prm%Rgas_inf = Rgas
      call sod2d_set_numerical_params(prm)
      call sod2d_hex64_tables(atoIJK, xgp, wgp, Ngp, dNgp, dlxigp_ip, invAtoIJK, gI, gJ, gK)
      ! the reference's "!$acc enter data copyin(...)" / "create(...)"
      call reg(c_loc(atoIJK), 64*i4, .true.); call reg(c_loc(invAtoIJK), 64*i4, .true.)
      call reg(c_loc(gI), 64*i4, .true.); call reg(c_loc(gJ), 64*i4, .true.); call reg(c_loc(gK), 64*i4, .true.)
      call reg(c_loc(xgp), 192*r8, .true.); call reg(c_loc(wgp), 64*r8, .true.)
      call reg(c_loc(Ngp), 4096*r8, .true.); call reg(c_loc(dNgp), 12288*r8, .true.)
      call reg(c_loc(dlxigp_ip), 768*r8, .true.)
      call reg(c_loc(connecOrig), 64*i4*nelem, .true.)
      call reg(c_loc(connec), 64*i4*nelem, .true.); call reg(c_loc(lpoin_w), i4*npoin_w, .true.)
      call reg(c_loc(point2elem), i4*npoin, .true.); call reg(c_loc(lnbn_nodes), i4*npoin, .true.)
      call reg(c_loc(coord), 3*r8*npoin, .true.)
      call reg(c_loc(rho), 3*r8*npoin, .true.); call reg(c_loc(uu), 9*r8*npoin, .true.)
      call reg(c_loc(q), 9*r8*npoin, .true.); call reg(c_loc(pr), 3*r8*npoin, .true.)
      call reg(c_loc(E), 3*r8*npoin, .true.); call reg(c_loc(Tem), 2*r8*npoin, .true.)
      call reg(c_loc(e_int), 2*r8*npoin, .true.); call reg(c_loc(eta), 3*r8*npoin, .true.)
      call reg(c_loc(csound), r8*npoin, .true.); call reg(c_loc(machno), r8*npoin, .true.)
      call reg(c_loc(mu_fluid), r8*npoin, .true.); call reg(c_loc(mu_factor), r8*npoin, .true.)
      call reg(c_loc(mu_e), 64*r8*nelem, .true.); call reg(c_loc(mu_sgs), 64*r8*nelem, .true.)
      call reg(c_loc(mue_l), 64*r8*nelem, .true.)
      call reg(c_loc(kres), r8*npoin, .true.); call reg(c_loc(etot), r8*npoin, .true.)
      call reg(c_loc(au), 3*r8*npoin, .true.); call reg(c_loc(ax1), r8*npoin, .true.)
      call reg(c_loc(ax2), r8*npoin, .true.); call reg(c_loc(ax3), r8*npoin, .true.)
      call reg(c_loc(fq), 3*r8*npoin, .true.)
      call reg(c_loc(gpvol), 64*r8*nelem, .false.); call reg(c_loc(He), 576*r8*nelem, .false.)
      call reg(c_loc(Ml), r8*npoin, .false.); call reg(c_loc(helem), r8*nelem, .false.)
      call reg(c_loc(helem_l), 64*r8*nelem, .false.)
      call reg(c_loc(Rmass), r8*npoin, .false.); call reg(c_loc(Rmom), 3*r8*npoin, .false.)
      call reg(c_loc(Rener), r8*npoin, .false.); call reg(c_loc(Rconv), r8*npoin, .false.)
      call reg(c_loc(Dmass), r8*npoin, .false.); call reg(c_loc(Dmom), 3*r8*npoin, .false.)
      call reg(c_loc(Dener), r8*npoin, .false.)

      ! initialisation as CFDSolverBase does it: geometry on the original
      ! connectivity (periodic copies keep their own coordinates), the rest
      ! on the working connectivity
      call elem_jacobian(nelem, npoin, connecOrig, coord, dNgp, wgp, gpvol, He)
      call lumped_mass_spectral(nelem, npoin, connec, gpvol, Ml)
      call sod2d_char_length(nelem, npoin, connecOrig, coord, helem)
      call sod2d_char_length_spectral(nelem, npoin, connecOrig, coord, helem_l)

      ! one call of each element routine on the level-2 state
      call full_convec_ijk(nelem, npoin, connec, Ngp, dNgp, He, gpvol, dlxigp_ip, xgp, invAtoIJK, gI, gJ, gK, &
                           uu(:, :, 2), q(:, :, 2), rho(:, 2), pr(:, 2), E(:, 2), Rmass, Rmom, Rener)
      call generic_scalar_convec_ijk(nelem, npoin, connec, Ngp, dNgp, He, gpvol, dlxigp_ip, xgp, invAtoIJK, &
                                     gI, gJ, gK, fq, eta(:, 2), uu(:, :, 2), Rconv)
      call full_diffusion_ijk(nelem, npoin, connec, Ngp, He, gpvol, dlxigp_ip, invAtoIJK, gI, gJ, gK, Cp, Prt, &
                              rho(:, 2), rho(:, 2), uu(:, :, 2), Tem(:, 2), mu_fluid, mu_e, mu_sgs, Ml, &
                              Dmass, Dmom, Dener)

      ! the time loop: dt from the CFL condition, then one RK4 step
      cfl = 0.5d0
      dt_cfl = 1.0d30
      call adapt_dt_cfl(nelem, npoin, connec, helem, uu(:, :, 2), csound, cfl, dt_cfl, cfl, mu_fluid, mu_sgs, &
                        rho(:, 2))
      call rk_4_main(.true., .false., nelem, 0, npoin, npoin_w, 0, point2elem, lnbn_nodes, dlxigp_ip, xgp, &
                     atoIJK, invAtoIJK, gI, gJ, gK, 1, connec, Ngp, dNgp, coord, wgp, He, Ml, gpvol, dt, &
                     helem, helem_l, Rgas, gamma_gas, Cp, Prt, rho, uu, q, pr, E, Tem, csound, machno, e_int, &
                     eta, mu_e, mu_sgs, kres, etot, au, ax1, ax2, ax3, lpoin_w, mu_fluid, mu_factor, mue_l)
      call volAvg_EK(nelem, npoin, connec, gpvol, Ngp, 1.0d0, rho(:, 2), uu(:, :, 2), EK)

      ! "!$acc update host(...)"
      call sod2d_update_host(c_loc(gpvol), 64*r8*nelem); call sod2d_update_host(c_loc(He), 576*r8*nelem)
      call sod2d_update_host(c_loc(Ml), r8*npoin); call sod2d_update_host(c_loc(helem), r8*nelem)
      call sod2d_update_host(c_loc(helem_l), 64*r8*nelem)
      call sod2d_update_host(c_loc(Rmass), r8*npoin); call sod2d_update_host(c_loc(Rmom), 3*r8*npoin)
      call sod2d_update_host(c_loc(Rener), r8*npoin); call sod2d_update_host(c_loc(Rconv), r8*npoin)
      call sod2d_update_host(c_loc(Dmass), r8*npoin); call sod2d_update_host(c_loc(Dmom), 3*r8*npoin)
      call sod2d_update_host(c_loc(Dener), r8*npoin)
      call sod2d_update_host(c_loc(rho), 3*r8*npoin); call sod2d_update_host(c_loc(q), 9*r8*npoin)
      call sod2d_update_host(c_loc(E), 3*r8*npoin); call sod2d_update_host(c_loc(mu_e), 64*r8*nelem)

      open (newunit=u, file=fout, access='stream', form='unformatted', status='replace')
      write (u) gpvol, He, Ml, helem, helem_l, Rmass, Rmom, Rener, Rconv, Dmass, Dmom, Dener
      write (u) dt_cfl, EK, rho(:, 2), q(:, :, 2), E(:, 2), mu_e
      close (u)
      call sod2d_finalize()
   end subroutine

end program shim_driver
