! mod_sod2d_hip -- ISO_C_BINDING drop-in for SOD2D's explicit compressible
! right-hand side + RK4 step on one MI355X (libsod2d_hip.so, p = 3, FP64).
!
! Every public routine below keeps the name, dummy list, order, kinds,
! extents and intents of the reference routine it replaces (file:line
! cited per routine, relative to src/lib_sod2d/sources/), so a caller swaps
!     use elem_convec, only: full_convec_ijk
! for
!     use mod_sod2d_hip, only: full_convec_ijk
! and keeps its call sites.  The routines forward to the C ABI of
! include/sod2d_hip.h; a non-zero return becomes ERROR STOP with the
! library's message (the reference routines have no error returns either).
!
! Array arguments are resolved by the library: device addresses are used as
! they are; host arrays must be registered first with sod2d_enter_data (the
! counterpart of the reference's "!$acc enter data copyin/create", e.g.
! time_integ.f90:37-68), which keeps a device copy that the hot-path routines
! read and write; sod2d_update_host / sod2d_update_device move data as
! "!$acc update host/device" does.
module mod_sod2d_hip
   use iso_c_binding
   use iso_fortran_env, only: error_unit
   implicit none
   private

   integer, parameter :: rp = c_double
   integer(4), parameter :: porder = 3, nnode = 64, ngaus = 64, ndime = 3, npbou = 16

   public :: sod2d_init, sod2d_finalize, sod2d_set_sync, sod2d_synchronize, sod2d_set_rk4_path
   public :: sod2d_enter_data, sod2d_exit_data, sod2d_update_device, sod2d_update_host
   public :: sod2d_numerical_params, sod2d_get_numerical_params, sod2d_set_numerical_params
   public :: sod2d_hex64_tables, sod2d_structured_box_mesh, sod2d_char_length, sod2d_char_length_spectral
   public :: elem_jacobian, lumped_mass_spectral
   public :: full_convec_ijk, full_convec_ijk_H, generic_scalar_convec_ijk, full_diffusion_ijk
   public :: sutherland_viscosity, lumped_solver_scal, lumped_solver_vect, smart_visc_spectral
   public :: adapt_dt_cfl, volAvg_EK, rk_4_main, sod2d_full_rhs

   ! mod_numerical_params.f90 module variables read by the path (the C
   ! struct sod_numerical_params)
   type, bind(C) :: sod2d_numerical_params
      real(c_double) :: ce, cmax, cglob, c_rho, c_ener, Rgas_inf
      integer(c_int) :: flag_normalise_entropy, flag_total_enthalpy, flag_rk_order
      integer(c_int) :: flag_les, flag_real_diff, flag_diff_suth
   end type

   interface
      integer(c_int) function c_init(p) bind(C, name="sod_init")
         import :: c_int
         integer(c_int), value :: p
      end function
      integer(c_int) function c_finalize() bind(C, name="sod_finalize")
         import :: c_int
      end function
      type(c_ptr) function c_last_error() bind(C, name="sod_last_error")
         import :: c_ptr
      end function
      integer(c_int) function c_set_sync(on) bind(C, name="sod_set_sync")
         import :: c_int
         integer(c_int), value :: on
      end function
      integer(c_int) function c_synchronize() bind(C, name="sod_synchronize")
         import :: c_int
      end function
      integer(c_int) function c_set_rk4_path(path) bind(C, name="sod_set_rk4_path")
         import :: c_int
         integer(c_int), value :: path
      end function
      integer(c_int) function c_get_params(p) bind(C, name="sod_get_numerical_params")
         import :: c_int, sod2d_numerical_params
         type(sod2d_numerical_params), intent(out) :: p
      end function
      integer(c_int) function c_set_params(p) bind(C, name="sod_set_numerical_params")
         import :: c_int, sod2d_numerical_params
         type(sod2d_numerical_params), intent(in) :: p
      end function
      integer(c_int) function c_enter_data(host, bytes, copyin) bind(C, name="sod_enter_data")
         import :: c_int, c_ptr, c_size_t
         type(c_ptr), value :: host
         integer(c_size_t), value :: bytes
         integer(c_int), value :: copyin
      end function
      integer(c_int) function c_exit_data(host, copyout) bind(C, name="sod_exit_data")
         import :: c_int, c_ptr
         type(c_ptr), value :: host
         integer(c_int), value :: copyout
      end function
      integer(c_int) function c_update_device(host, bytes) bind(C, name="sod_update_device")
         import :: c_int, c_ptr, c_size_t
         type(c_ptr), value :: host
         integer(c_size_t), value :: bytes
      end function
      integer(c_int) function c_update_host(host, bytes) bind(C, name="sod_update_host")
         import :: c_int, c_ptr, c_size_t
         type(c_ptr), value :: host
         integer(c_size_t), value :: bytes
      end function
      integer(c_int) function c_hex64_tables(atoIJK, xgp, wgp, Ngp, dNgp, dlxigp_ip, invAtoIJK, &
                                             gmshAtoI, gmshAtoJ, gmshAtoK) bind(C, name="sod_hex64_tables")
         import :: c_int, c_ptr
         type(c_ptr), value :: atoIJK, xgp, wgp, Ngp, dNgp, dlxigp_ip, invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK
      end function
      integer(c_int) function c_box_mesh(nx, ny, nz, Lx, Ly, Lz, amp, order, connec_orig, connec, lpoin_w, &
                                         coord, lattice) bind(C, name="sod_structured_box_mesh")
         import :: c_int, c_double, c_ptr
         integer(c_int), value :: nx, ny, nz
         real(c_double), value :: Lx, Ly, Lz, amp
         type(c_ptr), value :: order, connec_orig, connec, lpoin_w, coord, lattice
      end function
      integer(c_int) function c_elem_jacobian(nelem, npoin, connec, coord, dNgp, wgp, gpvol, He) &
            bind(C, name="sod_elem_jacobian")
         import :: c_int, c_ptr
         integer(c_int), value :: nelem, npoin
         type(c_ptr), value :: connec, coord, dNgp, wgp, gpvol, He
      end function
      integer(c_int) function c_lumped_mass(nelem, npoin, connec, gpvol, Ml) bind(C, name="sod_lumped_mass_spectral")
         import :: c_int, c_ptr
         integer(c_int), value :: nelem, npoin
         type(c_ptr), value :: connec, gpvol, Ml
      end function
      integer(c_int) function c_char_length(nelem, npoin, connec, coord, he) bind(C, name="sod_char_length")
         import :: c_int, c_ptr
         integer(c_int), value :: nelem, npoin
         type(c_ptr), value :: connec, coord, he
      end function
      integer(c_int) function c_char_length_spectral(nelem, npoin, connec, coord, he_l) &
            bind(C, name="sod_char_length_spectral")
         import :: c_int, c_ptr
         integer(c_int), value :: nelem, npoin
         type(c_ptr), value :: connec, coord, he_l
      end function
      integer(c_int) function c_full_convec(nelem, npoin, connec, Ngp, dNgp, He, gpvol, dlxigp_ip, xgp, &
                                            invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK, u, q, rho, pr, E, &
                                            Rmass, Rmom, Rener) bind(C, name="sod_full_convec_ijk")
         import :: c_int, c_ptr
         integer(c_int), value :: nelem, npoin
         type(c_ptr), value :: connec, Ngp, dNgp, He, gpvol, dlxigp_ip, xgp, invAtoIJK, gmshAtoI, gmshAtoJ, &
                               gmshAtoK, u, q, rho, pr, E, Rmass, Rmom, Rener
      end function
      integer(c_int) function c_full_convec_H(nelem, npoin, connec, Ngp, dNgp, He, gpvol, dlxigp_ip, xgp, &
                                              invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK, u, q, rho, pr, E, &
                                              Rmass, Rmom, Rener) bind(C, name="sod_full_convec_ijk_H")
         import :: c_int, c_ptr
         integer(c_int), value :: nelem, npoin
         type(c_ptr), value :: connec, Ngp, dNgp, He, gpvol, dlxigp_ip, xgp, invAtoIJK, gmshAtoI, gmshAtoJ, &
                               gmshAtoK, u, q, rho, pr, E, Rmass, Rmom, Rener
      end function
      integer(c_int) function c_scalar_convec(nelem, npoin, connec, Ngp, dNgp, He, gpvol, dlxigp_ip, xgp, &
                                              invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK, q, eta, u, Rconvec) &
            bind(C, name="sod_generic_scalar_convec_ijk")
         import :: c_int, c_ptr
         integer(c_int), value :: nelem, npoin
         type(c_ptr), value :: connec, Ngp, dNgp, He, gpvol, dlxigp_ip, xgp, invAtoIJK, gmshAtoI, gmshAtoJ, &
                               gmshAtoK, q, eta, u, Rconvec
      end function
      integer(c_int) function c_full_diffusion(nelem, npoin, connec, Ngp, He, gpvol, dlxigp_ip, invAtoIJK, &
                                               gmshAtoI, gmshAtoJ, gmshAtoK, Cp, Pr, rho_n, rho, u, Tem, &
                                               mu_fluid, mu_e, mu_sgs, Ml, Rmass, Rmom, Rener) &
            bind(C, name="sod_full_diffusion_ijk")
         import :: c_int, c_ptr, c_double
         integer(c_int), value :: nelem, npoin
         type(c_ptr), value :: connec, Ngp, He, gpvol, dlxigp_ip, invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK
         real(c_double), value :: Cp, Pr
         type(c_ptr), value :: rho_n, rho, u, Tem, mu_fluid, mu_e, mu_sgs, Ml, Rmass, Rmom, Rener
      end function
      integer(c_int) function c_sutherland(npoin, Tem, mu_factor, mu_fluid) bind(C, name="sod_sutherland_viscosity")
         import :: c_int, c_ptr
         integer(c_int), value :: npoin
         type(c_ptr), value :: Tem, mu_factor, mu_fluid
      end function
      integer(c_int) function c_lumped_scal(npoin, npoin_w, lpoin_w, Ml, R) bind(C, name="sod_lumped_solver_scal")
         import :: c_int, c_ptr
         integer(c_int), value :: npoin, npoin_w
         type(c_ptr), value :: lpoin_w, Ml, R
      end function
      integer(c_int) function c_lumped_vect(npoin, npoin_w, lpoin_w, Ml, R) bind(C, name="sod_lumped_solver_vect")
         import :: c_int, c_ptr
         integer(c_int), value :: npoin, npoin_w
         type(c_ptr), value :: lpoin_w, Ml, R
      end function
      integer(c_int) function c_smart_visc(nelem, npoin, npoin_w, connec, lpoin_w, Reta, Ngp, coord, dNgp, &
                                           gpvol, wgp, gamma_gas, rho, u, csound, Tem, eta, helem, helem_k, &
                                           Ml, mu_e, invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK, mue_l) &
            bind(C, name="sod_smart_visc_spectral")
         import :: c_int, c_ptr, c_double
         integer(c_int), value :: nelem, npoin, npoin_w
         type(c_ptr), value :: connec, lpoin_w, Reta, Ngp, coord, dNgp, gpvol, wgp
         real(c_double), value :: gamma_gas
         type(c_ptr), value :: rho, u, csound, Tem, eta, helem, helem_k, Ml, mu_e, invAtoIJK, gmshAtoI, &
                               gmshAtoJ, gmshAtoK, mue_l
      end function
      integer(c_int) function c_adapt_dt(nelem, npoin, connec, helem, u, csound, cfl_conv, dt, cfl_diff, &
                                         mu_fluid, mu_sgs, rho) bind(C, name="sod_adapt_dt_cfl")
         import :: c_int, c_ptr, c_double
         integer(c_int), value :: nelem, npoin
         type(c_ptr), value :: connec, helem, u, csound
         real(c_double), value :: cfl_conv
         type(c_ptr), value :: dt, cfl_diff, mu_fluid, mu_sgs, rho
      end function
      integer(c_int) function c_volavg_ek(nelem, npoin, connec, gpvol, Ngp, rho0, rho, u, EK) &
            bind(C, name="sod_volAvg_EK")
         import :: c_int, c_ptr, c_double
         integer(c_int), value :: nelem, npoin
         type(c_ptr), value :: connec, gpvol, Ngp
         real(c_double), value :: rho0
         type(c_ptr), value :: rho, u, EK
      end function
      integer(c_int) function c_full_rhs(nelem, npoin, npoin_w, connec, lpoin_w, He, gpvol, Ml, rho, q, E, &
                                         mu_factor, mu_e, mu_sgs, Rgas, gamma_gas, Cp, Prt, Rmass, Rmom, &
                                         Rener, Reta) bind(C, name="sod_full_rhs")
         import :: c_int, c_ptr, c_double
         integer(c_int), value :: nelem, npoin, npoin_w
         type(c_ptr), value :: connec, lpoin_w, He, gpvol, Ml, rho, q, E, mu_factor, mu_e, mu_sgs
         real(c_double), value :: Rgas, gamma_gas, Cp, Prt
         type(c_ptr), value :: Rmass, Rmom, Rener, Reta
      end function
      integer(c_int) function c_rk_4_main(noBoundaries, isWallModelOn, nelem, nboun, npoin, npoin_w, &
                                          numBoundsWM, point2elem, lnbn_nodes, dlxigp_ip, xgp, atoIJK, &
                                          invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK, ppow, connec, Ngp, dNgp, &
                                          coord, wgp, He, Ml, gpvol, dt, helem, helem_l, Rgas, gamma_gas, Cp, &
                                          Prt, rho, u, q, pr, E, Tem, csound, machno, e_int, eta, mu_e, &
                                          mu_sgs, kres, etot, au, ax1, ax2, ax3, lpoin_w, mu_fluid, &
                                          mu_factor, mue_l) bind(C, name="sod_rk_4_main")
         import :: c_int, c_ptr, c_double
         integer(c_int), value :: noBoundaries, isWallModelOn, nelem, nboun, npoin, npoin_w, numBoundsWM
         type(c_ptr), value :: point2elem, lnbn_nodes, dlxigp_ip, xgp, atoIJK, invAtoIJK, gmshAtoI, &
                               gmshAtoJ, gmshAtoK
         integer(c_int), value :: ppow
         type(c_ptr), value :: connec, Ngp, dNgp, coord, wgp, He, Ml, gpvol
         real(c_double), value :: dt
         type(c_ptr), value :: helem, helem_l
         real(c_double), value :: Rgas, gamma_gas, Cp, Prt
         type(c_ptr), value :: rho, u, q, pr, E, Tem, csound, machno, e_int, eta, mu_e, mu_sgs, kres, etot, &
                               au, ax1, ax2, ax3, lpoin_w, mu_fluid, mu_factor, mue_l
      end function
   end interface

contains

   ! ------------------------------------------------------------ errors
   subroutine check(rc, what)
      integer(c_int), intent(in) :: rc
      character(*), intent(in) :: what
      character(kind=c_char), pointer :: msg(:)
      character(len=512) :: text
      type(c_ptr) :: p
      integer :: i
      if (rc == 0) return
      text = ''
      p = c_last_error()
      if (c_associated(p)) then
         call c_f_pointer(p, msg, [512])
         do i = 1, 512
            if (msg(i) == c_null_char) exit
            text(i:i) = msg(i)
         end do
      end if
      write (error_unit, '(a,a,a,i0,a,a)') 'libsod2d_hip: ', what, ' returned ', rc, ': ', trim(text)
      error stop 1
   end subroutine check

   ! ------------------------------------------------------------ runtime
   subroutine sod2d_init()
      call check(c_init(int(porder, c_int)), 'sod_init')
   end subroutine
   subroutine sod2d_finalize()
      call check(c_finalize(), 'sod_finalize')
   end subroutine
   subroutine sod2d_set_sync(on)
      logical, intent(in) :: on
      call check(c_set_sync(merge(1_c_int, 0_c_int, on)), 'sod_set_sync')
   end subroutine
   subroutine sod2d_synchronize()
      call check(c_synchronize(), 'sod_synchronize')
   end subroutine
   subroutine sod2d_set_rk4_path(path)
      integer(4), intent(in) :: path
      call check(c_set_rk4_path(int(path, c_int)), 'sod_set_rk4_path')
   end subroutine
   subroutine sod2d_get_numerical_params(p)
      type(sod2d_numerical_params), intent(out) :: p
      call check(c_get_params(p), 'sod_get_numerical_params')
   end subroutine
   subroutine sod2d_set_numerical_params(p)
      type(sod2d_numerical_params), intent(in) :: p
      call check(c_set_params(p), 'sod_set_numerical_params')
   end subroutine

   ! "!$acc enter data copyin(a)" (copyin = .true.) / "create(a)"
   subroutine sod2d_enter_data(a, bytes, copyin)
      type(c_ptr), intent(in) :: a
      integer(8), intent(in) :: bytes
      logical, intent(in) :: copyin
      call check(c_enter_data(a, int(bytes, c_size_t), merge(1_c_int, 0_c_int, copyin)), 'sod_enter_data')
   end subroutine
   ! "!$acc exit data copyout(a)" (copyout = .true.) / "delete(a)"
   subroutine sod2d_exit_data(a, copyout)
      type(c_ptr), intent(in) :: a
      logical, intent(in) :: copyout
      call check(c_exit_data(a, merge(1_c_int, 0_c_int, copyout)), 'sod_exit_data')
   end subroutine
   subroutine sod2d_update_device(a, bytes)
      type(c_ptr), intent(in) :: a
      integer(8), intent(in) :: bytes
      call check(c_update_device(a, int(bytes, c_size_t)), 'sod_update_device')
   end subroutine
   subroutine sod2d_update_host(a, bytes)
      type(c_ptr), intent(in) :: a
      integer(8), intent(in) :: bytes
      call check(c_update_host(a, int(bytes, c_size_t)), 'sod_update_host')
   end subroutine

   ! ------------------------------------------------- init-time tables
   ! GaussLobattoLegendre_hex (quadrature_rules.f90:51-78), hex_highorder
   ! (elem_hex.f90:22-43), atoIJKInverse (mod_geom.f90:123-144); host only
   subroutine sod2d_hex64_tables(atoIJK, xgp, wgp, Ngp, dNgp, dlxigp_ip, invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK)
      integer(4), intent(out), target :: atoIJK(nnode), invAtoIJK(porder+1, porder+1, porder+1)
      integer(4), intent(out), target :: gmshAtoI(nnode), gmshAtoJ(nnode), gmshAtoK(nnode)
      real(rp), intent(out), target :: xgp(ngaus, ndime), wgp(ngaus), Ngp(ngaus, nnode), dNgp(ndime, nnode, ngaus)
      real(rp), intent(out), target :: dlxigp_ip(ngaus, ndime, porder+1)
      call check(c_hex64_tables(c_loc(atoIJK), c_loc(xgp), c_loc(wgp), c_loc(Ngp), c_loc(dNgp), c_loc(dlxigp_ip), &
                                c_loc(invAtoIJK), c_loc(gmshAtoI), c_loc(gmshAtoJ), c_loc(gmshAtoK)), &
                 'sod_hex64_tables')
   end subroutine

   ! structured periodic box as the reference's mesh tool numbers it
   ! (mod_meshConversorTool.f90:2864-3076, 3944-4035); host only
   subroutine sod2d_structured_box_mesh(nx, ny, nz, Lx, Ly, Lz, amp, connecOrig, connec, lpoin_w, coord)
      integer(4), intent(in) :: nx, ny, nz
      real(rp), intent(in) :: Lx, Ly, Lz, amp
      integer(4), intent(out), target :: connecOrig(nx*ny*nz, nnode), connec(nx*ny*nz, nnode)
      integer(4), intent(out), target :: lpoin_w(27*nx*ny*nz)
      real(rp), intent(out), target :: coord((3*nx+1)*(3*ny+1)*(3*nz+1), ndime)
      call check(c_box_mesh(int(nx, c_int), int(ny, c_int), int(nz, c_int), Lx, Ly, Lz, amp, c_null_ptr, &
                            c_loc(connecOrig), c_loc(connec), c_loc(lpoin_w), c_loc(coord), c_null_ptr), &
                 'sod_structured_box_mesh')
   end subroutine

   ! char_length (mod_geom.f90:13-37) for every element at once
   subroutine sod2d_char_length(nelem, npoin, connec, coord, helem)
      integer(4), intent(in) :: nelem, npoin
      integer(4), intent(in), target :: connec(nelem, nnode)
      real(rp), intent(in), target :: coord(npoin, ndime)
      real(rp), intent(out), target :: helem(nelem)
      call check(c_char_length(nelem, npoin, c_loc(connec), c_loc(coord), c_loc(helem)), 'sod_char_length')
   end subroutine

   ! char_length_spectral (mod_geom.f90:39-66) for every element at once
   subroutine sod2d_char_length_spectral(nelem, npoin, connec, coord, helem_l)
      integer(4), intent(in) :: nelem, npoin
      integer(4), intent(in), target :: connec(nelem, nnode)
      real(rp), intent(in), target :: coord(npoin, ndime)
      real(rp), intent(out), target :: helem_l(nelem, nnode)
      call check(c_char_length_spectral(nelem, npoin, c_loc(connec), c_loc(coord), c_loc(helem_l)), &
                 'sod_char_length_spectral')
   end subroutine

   ! jacobian_oper.f90:9-138
   subroutine elem_jacobian(nelem, npoin, connec, coord, dNgp, wgp, gpvol, He)
      integer(4), intent(in) :: nelem, npoin
      integer(4), intent(in), target :: connec(nelem, nnode)
      real(rp), intent(in), target :: coord(npoin, ndime), dNgp(ndime, nnode, ngaus), wgp(ngaus)
      real(rp), intent(out), target :: gpvol(1, ngaus, nelem), He(ndime, ndime, ngaus, nelem)
      call check(c_elem_jacobian(nelem, npoin, c_loc(connec), c_loc(coord), c_loc(dNgp), c_loc(wgp), &
                                 c_loc(gpvol), c_loc(He)), 'sod_elem_jacobian')
   end subroutine

   ! mass_matrix.f90:11-42
   subroutine lumped_mass_spectral(nelem, npoin, connec, gpvol, Ml)
      integer(4), intent(in) :: nelem, npoin
      integer(4), intent(in), target :: connec(nelem, nnode)
      real(rp), intent(in), target :: gpvol(1, ngaus, nelem)
      real(rp), intent(out), target :: Ml(npoin)
      call check(c_lumped_mass(nelem, npoin, c_loc(connec), c_loc(gpvol), c_loc(Ml)), 'sod_lumped_mass_spectral')
   end subroutine

   ! ------------------------------------------------------------ hot path
   ! elem_convec.f90:211-224
   subroutine full_convec_ijk(nelem, npoin, connec, Ngp, dNgp, He, gpvol, dlxigp_ip, xgp, invAtoIJK, gmshAtoI, &
                              gmshAtoJ, gmshAtoK, u, q, rho, pr, E, Rmass, Rmom, Rener)
      integer(4), intent(in) :: nelem, npoin
      integer(4), intent(in), target :: connec(nelem, nnode)
      real(rp), intent(in), target :: Ngp(ngaus, nnode), dNgp(ndime, nnode, ngaus)
      real(rp), intent(in), target :: He(ndime, ndime, ngaus, nelem), xgp(ngaus, ndime), &
                                      dlxigp_ip(ngaus, ndime, porder+1)
      real(rp), intent(in), target :: gpvol(1, ngaus, nelem)
      integer(4), intent(in), target :: invAtoIJK(porder+1, porder+1, porder+1), gmshAtoI(nnode), &
                                        gmshAtoJ(nnode), gmshAtoK(nnode)
      real(rp), intent(in), target :: q(npoin, ndime), u(npoin, ndime), rho(npoin), pr(npoin), E(npoin)
      real(rp), intent(out), target :: Rmass(npoin), Rmom(npoin, ndime), Rener(npoin)
      call check(c_full_convec(nelem, npoin, c_loc(connec), c_loc(Ngp), c_loc(dNgp), c_loc(He), c_loc(gpvol), &
                               c_loc(dlxigp_ip), c_loc(xgp), c_loc(invAtoIJK), c_loc(gmshAtoI), c_loc(gmshAtoJ), &
                               c_loc(gmshAtoK), c_loc(u), c_loc(q), c_loc(rho), c_loc(pr), c_loc(E), &
                               c_loc(Rmass), c_loc(Rmom), c_loc(Rener)), 'sod_full_convec_ijk')
   end subroutine

   ! elem_convec.f90:20-33
   subroutine full_convec_ijk_H(nelem, npoin, connec, Ngp, dNgp, He, gpvol, dlxigp_ip, xgp, invAtoIJK, gmshAtoI, &
                                gmshAtoJ, gmshAtoK, u, q, rho, pr, E, Rmass, Rmom, Rener)
      integer(4), intent(in) :: nelem, npoin
      integer(4), intent(in), target :: connec(nelem, nnode)
      real(rp), intent(in), target :: Ngp(ngaus, nnode), dNgp(ndime, nnode, ngaus)
      real(rp), intent(in), target :: He(ndime, ndime, ngaus, nelem), xgp(ngaus, ndime), &
                                      dlxigp_ip(ngaus, ndime, porder+1)
      real(rp), intent(in), target :: gpvol(1, ngaus, nelem)
      integer(4), intent(in), target :: invAtoIJK(porder+1, porder+1, porder+1), gmshAtoI(nnode), &
                                        gmshAtoJ(nnode), gmshAtoK(nnode)
      real(rp), intent(in), target :: q(npoin, ndime), u(npoin, ndime), rho(npoin), pr(npoin), E(npoin)
      real(rp), intent(out), target :: Rmass(npoin), Rmom(npoin, ndime), Rener(npoin)
      call check(c_full_convec_H(nelem, npoin, c_loc(connec), c_loc(Ngp), c_loc(dNgp), c_loc(He), c_loc(gpvol), &
                                 c_loc(dlxigp_ip), c_loc(xgp), c_loc(invAtoIJK), c_loc(gmshAtoI), c_loc(gmshAtoJ), &
                                 c_loc(gmshAtoK), c_loc(u), c_loc(q), c_loc(rho), c_loc(pr), c_loc(E), &
                                 c_loc(Rmass), c_loc(Rmom), c_loc(Rener)), 'sod_full_convec_ijk_H')
   end subroutine

   ! elem_convec.f90:462-476
   subroutine generic_scalar_convec_ijk(nelem, npoin, connec, Ngp, dNgp, He, gpvol, dlxigp_ip, xgp, invAtoIJK, &
                                        gmshAtoI, gmshAtoJ, gmshAtoK, q, eta, u, Rconvec)
      integer(4), intent(in) :: nelem, npoin
      integer(4), intent(in), target :: connec(nelem, nnode)
      real(rp), intent(in), target :: Ngp(ngaus, nnode), dNgp(ndime, nnode, ngaus)
      real(rp), intent(in), target :: He(ndime, ndime, ngaus, nelem), xgp(ngaus, ndime), &
                                      dlxigp_ip(ngaus, ndime, porder+1)
      real(rp), intent(in), target :: gpvol(1, ngaus, nelem)
      integer(4), intent(in), target :: invAtoIJK(porder+1, porder+1, porder+1), gmshAtoI(nnode), &
                                        gmshAtoJ(nnode), gmshAtoK(nnode)
      real(rp), intent(in), target :: q(npoin, ndime), eta(npoin), u(npoin, ndime)
      real(rp), intent(out), target :: Rconvec(npoin)
      call check(c_scalar_convec(nelem, npoin, c_loc(connec), c_loc(Ngp), c_loc(dNgp), c_loc(He), c_loc(gpvol), &
                                 c_loc(dlxigp_ip), c_loc(xgp), c_loc(invAtoIJK), c_loc(gmshAtoI), c_loc(gmshAtoJ), &
                                 c_loc(gmshAtoK), c_loc(q), c_loc(eta), c_loc(u), c_loc(Rconvec)), &
                 'sod_generic_scalar_convec_ijk')
   end subroutine

   ! elem_diffu.f90:12-25
   subroutine full_diffusion_ijk(nelem, npoin, connec, Ngp, He, gpvol, dlxigp_ip, invAtoIJK, gmshAtoI, gmshAtoJ, &
                                 gmshAtoK, Cp, Pr, rho_n, rho, u, Tem, mu_fluid, mu_e, mu_sgs, Ml, Rmass, Rmom, Rener)
      integer(4), intent(in) :: nelem, npoin
      integer(4), intent(in), target :: connec(nelem, nnode)
      real(rp), intent(in), target :: Ngp(ngaus, nnode)
      real(rp), intent(in), target :: He(ndime, ndime, ngaus, nelem), dlxigp_ip(ngaus, ndime, porder+1)
      real(rp), intent(in), target :: gpvol(1, ngaus, nelem)
      integer(4), intent(in), target :: invAtoIJK(porder+1, porder+1, porder+1), gmshAtoI(nnode), &
                                        gmshAtoJ(nnode), gmshAtoK(nnode)
      real(rp), intent(in) :: Cp, Pr
      real(rp), intent(in), target :: rho_n(npoin), rho(npoin), u(npoin, ndime), Tem(npoin), &
                                      mu_e(nelem, ngaus), mu_sgs(nelem, ngaus), Ml(npoin)
      real(rp), intent(in), target :: mu_fluid(npoin)
      real(rp), intent(out), target :: Rmass(npoin), Rmom(npoin, ndime), Rener(npoin)
      call check(c_full_diffusion(nelem, npoin, c_loc(connec), c_loc(Ngp), c_loc(He), c_loc(gpvol), &
                                  c_loc(dlxigp_ip), c_loc(invAtoIJK), c_loc(gmshAtoI), c_loc(gmshAtoJ), &
                                  c_loc(gmshAtoK), Cp, Pr, c_loc(rho_n), c_loc(rho), c_loc(u), c_loc(Tem), &
                                  c_loc(mu_fluid), c_loc(mu_e), c_loc(mu_sgs), c_loc(Ml), c_loc(Rmass), &
                                  c_loc(Rmom), c_loc(Rener)), 'sod_full_diffusion_ijk')
   end subroutine

   ! mod_fluid_viscosity.f90:21-41
   subroutine sutherland_viscosity(npoin, Tem, mu_factor, mu_fluid)
      integer(4), intent(in) :: npoin
      real(rp), intent(in), target :: Tem(npoin), mu_factor(npoin)
      real(rp), intent(inout), target :: mu_fluid(npoin)
      call check(c_sutherland(npoin, c_loc(Tem), c_loc(mu_factor), c_loc(mu_fluid)), 'sod_sutherland_viscosity')
   end subroutine

   ! mod_solver.f90:23-38
   subroutine lumped_solver_scal(npoin, npoin_w, lpoin_w, Ml, R)
      integer(4), intent(in) :: npoin, npoin_w
      integer(4), intent(in), target :: lpoin_w(npoin_w)
      real(rp), intent(in), target :: Ml(npoin)
      real(rp), intent(inout), target :: R(npoin)
      call check(c_lumped_scal(npoin, npoin_w, c_loc(lpoin_w), c_loc(Ml), c_loc(R)), 'sod_lumped_solver_scal')
   end subroutine

   ! mod_solver.f90:40-58
   subroutine lumped_solver_vect(npoin, npoin_w, lpoin_w, Ml, R)
      integer(4), intent(in) :: npoin, npoin_w
      integer(4), intent(in), target :: lpoin_w(npoin_w)
      real(rp), intent(in), target :: Ml(npoin)
      real(rp), intent(inout), target :: R(npoin, ndime)
      call check(c_lumped_vect(npoin, npoin_w, c_loc(lpoin_w), c_loc(Ml), c_loc(R)), 'sod_lumped_solver_vect')
   end subroutine

   ! mod_entropy_viscosity.f90:14-27
   subroutine smart_visc_spectral(nelem, npoin, npoin_w, connec, lpoin_w, Reta, Ngp, coord, dNgp, gpvol, wgp, &
                                  gamma_gas, rho, u, csound, Tem, eta, helem, helem_k, Ml, mu_e, invAtoIJK, &
                                  gmshAtoI, gmshAtoJ, gmshAtoK, mue_l)
      integer(4), intent(in) :: nelem, npoin, npoin_w
      integer(4), intent(in), target :: connec(nelem, nnode), lpoin_w(npoin_w)
      real(rp), intent(in), target :: Reta(npoin), Ngp(ngaus, nnode)
      real(rp), intent(in) :: gamma_gas
      real(rp), intent(in), target :: rho(npoin), u(npoin, ndime), csound(npoin), Tem(npoin), eta(npoin), &
                                      helem(nelem, nnode), helem_k(nelem), Ml(npoin)
      real(rp), intent(out), target :: mu_e(nelem, ngaus)
      real(rp), intent(in), target :: coord(npoin, ndime), dNgp(ndime, nnode, ngaus), wgp(ngaus)
      real(rp), intent(in), target :: gpvol(1, ngaus, nelem)
      integer(4), intent(in), target :: invAtoIJK(porder+1, porder+1, porder+1), gmshAtoI(nnode), &
                                        gmshAtoJ(nnode), gmshAtoK(nnode)
      real(rp), intent(inout), target :: mue_l(nelem, nnode)
      call check(c_smart_visc(nelem, npoin, npoin_w, c_loc(connec), c_loc(lpoin_w), c_loc(Reta), c_loc(Ngp), &
                              c_loc(coord), c_loc(dNgp), c_loc(gpvol), c_loc(wgp), gamma_gas, c_loc(rho), c_loc(u), &
                              c_loc(csound), c_loc(Tem), c_loc(eta), c_loc(helem), c_loc(helem_k), c_loc(Ml), &
                              c_loc(mu_e), c_loc(invAtoIJK), c_loc(gmshAtoI), c_loc(gmshAtoJ), c_loc(gmshAtoK), &
                              c_loc(mue_l)), 'sod_smart_visc_spectral')
   end subroutine

   ! mod_time_ops.f90:11-61; dt is a host scalar
   subroutine adapt_dt_cfl(nelem, npoin, connec, helem, u, csound, cfl_conv, dt, cfl_diff, mu_fluid, mu_sgs, rho)
      integer(4), intent(in) :: nelem, npoin
      integer(4), intent(in), target :: connec(nelem, nnode)
      real(rp), intent(in), target :: helem(nelem)
      real(rp), intent(in), target :: u(npoin, ndime), csound(npoin)
      real(rp), intent(in) :: cfl_conv
      real(rp), intent(in), optional, target :: cfl_diff, mu_fluid(npoin), mu_sgs(nelem, nnode), rho(npoin)
      real(rp), intent(inout), target :: dt
      if (present(cfl_diff)) then
         if (.not. (present(mu_fluid) .and. present(mu_sgs) .and. present(rho))) then
            write (error_unit, '(a)') 'adapt_dt_cfl: cfl_diff needs mu_fluid, mu_sgs and rho'
            error stop 1
         end if
         call check(c_adapt_dt(nelem, npoin, c_loc(connec), c_loc(helem), c_loc(u), c_loc(csound), cfl_conv, &
                               c_loc(dt), c_loc(cfl_diff), c_loc(mu_fluid), c_loc(mu_sgs), c_loc(rho)), &
                    'sod_adapt_dt_cfl')
      else
         call check(c_adapt_dt(nelem, npoin, c_loc(connec), c_loc(helem), c_loc(u), c_loc(csound), cfl_conv, &
                               c_loc(dt), c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr), 'sod_adapt_dt_cfl')
      end if
   end subroutine

   ! mod_analysis.f90:10-44; EK is a host scalar
   subroutine volAvg_EK(nelem, npoin, connec, gpvol, Ngp, rho0, rho, u, EK)
      integer(4), intent(in) :: nelem, npoin
      integer(4), intent(in), target :: connec(nelem, nnode)
      real(rp), intent(in), target :: gpvol(1, ngaus, nelem), Ngp(ngaus, nnode)
      real(rp), intent(in) :: rho0
      real(rp), intent(in), target :: rho(npoin), u(npoin, ndime)
      real(rp), intent(out), target :: EK
      call check(c_volavg_ek(nelem, npoin, c_loc(connec), c_loc(gpvol), c_loc(Ngp), rho0, c_loc(rho), c_loc(u), &
                             c_loc(EK)), 'sod_volAvg_EK')
   end subroutine

   ! One assembled right-hand side of an RK stage (time_integ.f90:266-373);
   ! no single reference routine computes this sum.
   subroutine sod2d_full_rhs(nelem, npoin, npoin_w, connec, lpoin_w, He, gpvol, Ml, rho, q, E, mu_factor, mu_e, &
                             mu_sgs, Rgas, gamma_gas, Cp, Prt, Rmass, Rmom, Rener, Reta)
      integer(4), intent(in) :: nelem, npoin, npoin_w
      integer(4), intent(in), target :: connec(nelem, nnode), lpoin_w(npoin_w)
      real(rp), intent(in), target :: He(ndime, ndime, ngaus, nelem), gpvol(1, ngaus, nelem), Ml(npoin)
      real(rp), intent(in), target :: rho(npoin), q(npoin, ndime), E(npoin), mu_factor(npoin)
      real(rp), intent(in), target :: mu_e(nelem, ngaus), mu_sgs(nelem, ngaus)
      real(rp), intent(in) :: Rgas, gamma_gas, Cp, Prt
      real(rp), intent(out), target :: Rmass(npoin), Rmom(npoin, ndime), Rener(npoin), Reta(npoin)
      call check(c_full_rhs(nelem, npoin, npoin_w, c_loc(connec), c_loc(lpoin_w), c_loc(He), c_loc(gpvol), &
                            c_loc(Ml), c_loc(rho), c_loc(q), c_loc(E), c_loc(mu_factor), c_loc(mu_e), &
                            c_loc(mu_sgs), Rgas, gamma_gas, Cp, Prt, c_loc(Rmass), c_loc(Rmom), c_loc(Rener), &
                            c_loc(Reta)), 'sod_full_rhs')
   end subroutine

   ! time_integ.f90:137-191.  The periodic explicit branch (noBoundaries =
   ! .true., no wall model, no LES); the optional boundary arguments are
   ! accepted and, as in that branch of the reference, not used.
   subroutine rk_4_main(noBoundaries, isWallModelOn, nelem, nboun, npoin, npoin_w, numBoundsWM, point2elem, &
                        lnbn_nodes, dlxigp_ip, xgp, atoIJK, invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK, ppow, connec, &
                        Ngp, dNgp, coord, wgp, He, Ml, gpvol, dt, helem, helem_l, Rgas, gamma_gas, Cp, Prt, rho, u, &
                        q, pr, E, Tem, csound, machno, e_int, eta, mu_e, mu_sgs, kres, etot, au, ax1, ax2, ax3, &
                        lpoin_w, mu_fluid, mu_factor, mue_l, ndof, nbnodes, ldof, lbnodes, bound, bou_codes, &
                        bou_codes_nodes, listBoundsWM, wgp_b, bounorm, normalsAtNodes, u_buffer, tauw, source_term, &
                        walave_u, zo)
      logical, intent(in) :: noBoundaries, isWallModelOn
      integer(4), intent(in) :: nelem, nboun, npoin
      integer(4), intent(in), target :: connec(nelem, nnode), lpoin_w(npoin_w), point2elem(npoin), lnbn_nodes(npoin)
      integer(4), intent(in) :: npoin_w
      integer(4), intent(in), target :: atoIJK(nnode), invAtoIJK(porder+1, porder+1, porder+1), gmshAtoI(nnode), &
                                        gmshAtoJ(nnode), gmshAtoK(nnode)
      integer(4), intent(in) :: ppow
      real(rp), intent(in), target :: Ngp(ngaus, nnode), dNgp(ndime, nnode, ngaus), dlxigp_ip(ngaus, ndime, porder+1)
      real(rp), intent(in), target :: He(ndime, ndime, ngaus, nelem), xgp(ngaus, ndime)
      real(rp), intent(in), target :: gpvol(1, ngaus, nelem)
      real(rp), intent(in) :: dt
      real(rp), intent(in), target :: helem(nelem), helem_l(nelem, nnode), Ml(npoin), mu_factor(npoin)
      real(rp), intent(in) :: Rgas, gamma_gas, Cp, Prt
      real(rp), intent(inout), target :: rho(npoin, 3), u(npoin, ndime, 3), q(npoin, ndime, 3), pr(npoin, 3), &
                                         E(npoin, 3), Tem(npoin, 2), e_int(npoin, 2), eta(npoin, 3)
      real(rp), intent(inout), target :: mu_fluid(npoin), csound(npoin), machno(npoin)
      real(rp), intent(inout), target :: mu_e(nelem, ngaus), mu_sgs(nelem, ngaus)
      real(rp), intent(inout), target :: kres(npoin), etot(npoin), au(npoin, ndime), ax1(npoin), ax2(npoin), ax3(npoin)
      real(rp), intent(inout), target :: mue_l(nelem, nnode)
      real(rp), intent(in), target :: coord(npoin, ndime), wgp(ngaus)
      integer(4), intent(in) :: numBoundsWM
      integer(4), optional, intent(in) :: ndof, nbnodes, ldof(*), lbnodes(*)
      integer(4), optional, intent(in) :: bound(nboun, npbou), bou_codes(nboun), bou_codes_nodes(npoin)
      integer(4), optional, intent(in) :: listBoundsWM(*)
      real(rp), optional, intent(in) :: wgp_b(npbou), bounorm(nboun, ndime*npbou), normalsAtNodes(npoin, ndime)
      real(rp), optional, intent(in) :: u_buffer(npoin, ndime)
      real(rp), optional, intent(inout) :: tauw(npoin, ndime)
      real(rp), optional, intent(in) :: source_term(npoin, ndime)
      real(rp), optional, intent(in) :: walave_u(npoin, ndime)
      real(rp), optional, intent(in) :: zo(npoin)
      if (present(source_term)) then
         write (error_unit, '(a)') 'rk_4_main: source terms are outside the periodic explicit path of libsod2d_hip'
         error stop 1
      end if
      call check(c_rk_4_main(merge(1_c_int, 0_c_int, noBoundaries), merge(1_c_int, 0_c_int, isWallModelOn), &
                             nelem, nboun, npoin, npoin_w, numBoundsWM, c_loc(point2elem), c_loc(lnbn_nodes), &
                             c_loc(dlxigp_ip), c_loc(xgp), c_loc(atoIJK), c_loc(invAtoIJK), c_loc(gmshAtoI), &
                             c_loc(gmshAtoJ), c_loc(gmshAtoK), ppow, c_loc(connec), c_loc(Ngp), c_loc(dNgp), &
                             c_loc(coord), c_loc(wgp), c_loc(He), c_loc(Ml), c_loc(gpvol), dt, c_loc(helem), &
                             c_loc(helem_l), Rgas, gamma_gas, Cp, Prt, c_loc(rho), c_loc(u), c_loc(q), c_loc(pr), &
                             c_loc(E), c_loc(Tem), c_loc(csound), c_loc(machno), c_loc(e_int), c_loc(eta), &
                             c_loc(mu_e), c_loc(mu_sgs), c_loc(kres), c_loc(etot), c_loc(au), c_loc(ax1), &
                             c_loc(ax2), c_loc(ax3), c_loc(lpoin_w), c_loc(mu_fluid), c_loc(mu_factor), &
                             c_loc(mue_l)), 'sod_rk_4_main')
   end subroutine

end module mod_sod2d_hip
