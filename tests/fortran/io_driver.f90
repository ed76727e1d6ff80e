! A Fortran caller of mod_sod2d_io: loads one partition of a SOD2D mesh file
! and a restart file, dumps what it read (stream file, for the test to
! compare), and writes a new restart from the loaded fields.
!   io_driver <mesh.hdf> <rank> <restart_in.h5> <restart_out.h5> <dump.bin>
program io_driver
   use iso_c_binding
   use mod_sod2d_io
   implicit none
   type(sod2d_mesh_partition) :: m
   character(len=1024) :: mesh_path, rst_in, rst_out, dump, arg
   integer(4) :: rank, istep, u_
   real(8) :: time
   real(8), allocatable :: rho(:), u(:,:), pr(:), E(:), mue(:,:), mut(:,:)
   call get_command_argument(1, mesh_path)
   call get_command_argument(2, arg)
   read(arg, *) rank
   call get_command_argument(3, rst_in)
   call get_command_argument(4, rst_out)
   call get_command_argument(5, dump)
   call sod2d_load_mesh(mesh_path, rank, m)
   allocate(rho(m%numNodesRankPar), u(m%numNodesRankPar, 3), pr(m%numNodesRankPar), E(m%numNodesRankPar))
   allocate(mue(m%numElemsRankPar, 64), mut(m%numElemsRankPar, 64))
   call sod2d_load_restart(rst_in, m%rankNodeStart, m%connecParOrig, m%masSlaRankPar, m%nPerRankPar, istep, time, &
                           rho, u, pr, E, mue, mut)
   print '(5i12,l3)', m%numElemsRankPar, m%numNodesRankPar, m%numWorkingNodesRankPar, m%nPerRankPar, istep, &
                      m%isMeshPeriodic
   open(newunit=u_, file=trim(dump), access='stream', form='unformatted', status='replace')
   write(u_) m%connecParOrig, m%connecParWork, m%workingNodesPar, m%coordPar
   if (m%nPerRankPar > 0) write(u_) m%masSlaRankPar(1:m%nPerRankPar, :)
   write(u_) rho, u, pr, E, mue, mut, time
   close(u_)
   ! a restart of this partition alone: doubled density, next step
   call sod2d_save_restart(rst_out, .true., int(m%numNodesRankPar, 8), 1_8, m%connecParOrig, istep + 1, &
                           time + 1.0d0, 2.0d0 * rho, u, pr, E, mue, mut)
end program io_driver
