! mod_sod2d_io -- ISO_C_BINDING access to libsod2d_io.so (include/sod2d_io.h):
! SOD2D's partitioned-HDF5 mesh files and restart files, read and written on
! the host.
!
! sod2d_load_mesh stands in for load_hdf5_meshFile
! (src/lib_sod2d/sources/mod_hdf5.f90:1294-1401) for one partition: it
! allocates and fills the arrays the reference keeps in mod_mpi_mesh
! (connecParOrig, connecParWork, workingNodesPar, coordPar, masSlaRankPar and
! the partition sizes).  sod2d_save_restart / sod2d_load_restart stand in for
! save_hdf5_restartFile / load_hdf5_restartFile (mod_hdf5.f90:3214-3436)
! with the reference's dummy arrays.  A non-zero return of the library
! becomes ERROR STOP with its message.
module mod_sod2d_io
   use iso_c_binding
   use iso_fortran_env, only: error_unit
   implicit none
   private

   integer, parameter :: rp = c_double
   integer(4), parameter :: nnode = 64

   public :: sod2d_mesh_partition, sod2d_load_mesh, sod2d_save_restart, sod2d_load_restart

   ! The arrays and sizes of one partition (mod_mpi_mesh's names).
   type :: sod2d_mesh_partition
      integer(4) :: nranks = 0, porder = 0
      integer(8) :: totalNumNodesPar = 0
      integer(4) :: totalNumElements = 0
      integer(4) :: numNodesRankPar = 0, numElemsRankPar = 0, numWorkingNodesRankPar = 0, nPerRankPar = 0
      integer(8) :: rankNodeStart = 0, rankNodeEnd = 0
      integer(4) :: rankElemStart = 0, rankElemEnd = 0
      logical :: isMeshPeriodic = .false.
      integer(4), allocatable :: connecParOrig(:,:), connecParWork(:,:), workingNodesPar(:)
      integer(4), allocatable :: masSlaRankPar(:,:)
      real(rp), allocatable :: coordPar(:,:)
   end type

   ! indices of the info vector (include/sod2d_io.h, 1-based here)
   integer, parameter :: I_NRANKS = 1, I_PORDER = 2, I_TOTAL_NODES = 4, I_TOTAL_ELEMS = 5, I_PERIODIC = 6
   integer, parameter :: I_NODE_START = 9, I_NODE_END = 10, I_ELEM_START = 11, I_ELEM_END = 12
   integer, parameter :: I_NPOIN = 13, I_NELEM = 14, I_NPOIN_W = 15, I_NPER = 16, I_COUNT = 18

   interface
      type(c_ptr) function c_io_last_error() bind(C, name="sod_io_last_error")
         import :: c_ptr
      end function
      integer(c_int) function c_mesh_open(path, rank, h) bind(C, name="sod_h5_mesh_open")
         import :: c_int, c_char, c_ptr
         character(kind=c_char), intent(in) :: path(*)
         integer(c_int), value :: rank
         type(c_ptr), intent(out) :: h
      end function
      integer(c_int) function c_mesh_info(h, info) bind(C, name="sod_h5_mesh_info")
         import :: c_int, c_ptr, c_long_long
         type(c_ptr), value :: h
         integer(c_long_long), intent(out) :: info(*)
      end function
      integer(c_int) function c_mesh_read(h, co, cw, wn, xyz, ms, gs, gp, eg, a2ijk, g2ijk, v2ijk) &
            bind(C, name="sod_h5_mesh_read")
         import :: c_int, c_ptr
         type(c_ptr), value :: h, co, cw, wn, xyz, ms, gs, gp, eg, a2ijk, g2ijk, v2ijk
      end function
      subroutine c_mesh_close(h) bind(C, name="sod_h5_mesh_close")
         import :: c_ptr
         type(c_ptr), value :: h
      end subroutine
      integer(c_int) function c_restart_save(path, create, total, start, npoin, nelem, co, istep, time, rho, u, pr, &
            E, mue, mut, walave) bind(C, name="sod_h5_restart_save")
         import :: c_int, c_char, c_long_long, c_double, c_ptr
         character(kind=c_char), intent(in) :: path(*)
         integer(c_int), value :: create, npoin, nelem, istep
         integer(c_long_long), value :: total, start
         real(c_double), value :: time
         type(c_ptr), value :: co, rho, u, pr, E, mue, mut, walave
      end function
      integer(c_int) function c_restart_load(path, start, npoin, nelem, co, nper, ms, istep, time, rho, u, pr, E, &
            mue, mut, walave) bind(C, name="sod_h5_restart_load")
         import :: c_int, c_char, c_long_long, c_double, c_ptr
         character(kind=c_char), intent(in) :: path(*)
         integer(c_long_long), value :: start
         integer(c_int), value :: npoin, nelem, nper
         type(c_ptr), value :: co, ms, rho, u, pr, E, mue, mut, walave
         integer(c_int), intent(out) :: istep
         real(c_double), intent(out) :: time
      end function
   end interface

contains

   subroutine check(rc, what)
      integer(c_int), intent(in) :: rc
      character(*), intent(in) :: what
      character(kind=c_char), pointer :: msg(:)
      integer :: n
      if (rc == 0) return
      call c_f_pointer(c_io_last_error(), msg, [1024])
      n = 0
      do while (n < 1024)
         if (msg(n + 1) == c_null_char) exit
         n = n + 1
      end do
      write(error_unit, '(a,": ",1024a)') what, msg(1:n)
      error stop 1
   end subroutine check

   ! load_hdf5_meshFile for partition `rank` of the file `path`.
   subroutine sod2d_load_mesh(path, rank, m)
      character(*), intent(in) :: path
      integer(4), intent(in) :: rank
      type(sod2d_mesh_partition), intent(out), target :: m
      type(c_ptr) :: h
      integer(c_long_long) :: info(I_COUNT)
      type(c_ptr) :: pms
      call check(c_mesh_open(trim(path)//c_null_char, int(rank, c_int), h), "sod2d_load_mesh: "//trim(path))
      call check(c_mesh_info(h, info), "sod2d_load_mesh")
      m%nranks = int(info(I_NRANKS), 4)
      m%porder = int(info(I_PORDER), 4)
      m%totalNumNodesPar = info(I_TOTAL_NODES)
      m%totalNumElements = int(info(I_TOTAL_ELEMS), 4)
      m%isMeshPeriodic = info(I_PERIODIC) /= 0
      m%rankNodeStart = info(I_NODE_START)
      m%rankNodeEnd = info(I_NODE_END)
      m%rankElemStart = int(info(I_ELEM_START), 4)
      m%rankElemEnd = int(info(I_ELEM_END), 4)
      m%numNodesRankPar = int(info(I_NPOIN), 4)
      m%numElemsRankPar = int(info(I_NELEM), 4)
      m%numWorkingNodesRankPar = int(info(I_NPOIN_W), 4)
      m%nPerRankPar = int(info(I_NPER), 4)
      allocate(m%connecParOrig(m%numElemsRankPar, nnode), m%connecParWork(m%numElemsRankPar, nnode))
      allocate(m%workingNodesPar(m%numWorkingNodesRankPar), m%coordPar(m%numNodesRankPar, 3))
      allocate(m%masSlaRankPar(max(m%nPerRankPar, 1), 2))
      pms = c_null_ptr
      if (m%nPerRankPar > 0) pms = c_loc(m%masSlaRankPar)
      ! masSlaRankPar is read as (nPer,2) column-major; with nPer = 0 nothing is read
      call check(c_mesh_read(h, c_loc(m%connecParOrig), c_loc(m%connecParWork), c_loc(m%workingNodesPar), &
                             c_loc(m%coordPar), pms, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, &
                             c_null_ptr), "sod2d_load_mesh")
      call c_mesh_close(h)
   end subroutine sod2d_load_mesh

   ! save_hdf5_restartFile for one partition (create = .true. makes the
   ! file and writes time and istep).
   subroutine sod2d_save_restart(path, create, totalNumNodesPar, rankNodeStart, connecParOrig, istep, time, &
                                 rho, u, pr, E, mu_e, mu_t)
      character(*), intent(in) :: path
      logical, intent(in) :: create
      integer(8), intent(in) :: totalNumNodesPar, rankNodeStart
      integer(4), intent(in), target, contiguous :: connecParOrig(:,:)
      integer(4), intent(in) :: istep
      real(rp), intent(in) :: time
      real(rp), intent(in), target, contiguous :: rho(:), u(:,:), pr(:), E(:), mu_e(:,:), mu_t(:,:)
      integer(c_int) :: cr
      cr = 0
      if (create) cr = 1
      call check(c_restart_save(trim(path)//c_null_char, cr, int(totalNumNodesPar, c_long_long), &
                                int(rankNodeStart, c_long_long), int(size(rho), c_int), &
                                int(size(connecParOrig, 1), c_int), c_loc(connecParOrig), int(istep, c_int), time, &
                                c_loc(rho), c_loc(u), c_loc(pr), c_loc(E), c_loc(mu_e), c_loc(mu_t), c_null_ptr), &
                 "sod2d_save_restart: "//trim(path))
   end subroutine sod2d_save_restart

   ! load_hdf5_restartFile for one partition.
   subroutine sod2d_load_restart(path, rankNodeStart, connecParOrig, masSlaRankPar, nPerRankPar, istep, time, &
                                 rho, u, pr, E, mu_e, mu_t)
      character(*), intent(in) :: path
      integer(8), intent(in) :: rankNodeStart
      integer(4), intent(in), target, contiguous :: connecParOrig(:,:), masSlaRankPar(:,:)
      integer(4), intent(in) :: nPerRankPar
      integer(4), intent(out) :: istep
      real(rp), intent(out) :: time
      real(rp), intent(inout), target, contiguous :: rho(:), u(:,:), pr(:), E(:), mu_e(:,:), mu_t(:,:)
      integer(c_int) :: cstep
      real(c_double) :: ctime
      type(c_ptr) :: pms
      pms = c_null_ptr
      if (nPerRankPar > 0) pms = c_loc(masSlaRankPar)
      call check(c_restart_load(trim(path)//c_null_char, int(rankNodeStart, c_long_long), int(size(rho), c_int), &
                                int(size(connecParOrig, 1), c_int), c_loc(connecParOrig), int(nPerRankPar, c_int), &
                                pms, cstep, ctime, c_loc(rho), c_loc(u), c_loc(pr), c_loc(E), c_loc(mu_e), &
                                c_loc(mu_t), c_null_ptr), "sod2d_load_restart: "//trim(path))
      istep = int(cstep, 4)
      time = real(ctime, rp)
   end subroutine sod2d_load_restart

end module mod_sod2d_io
