program partition_frame_driver
  ! Reads one 2-D frame from a binary stream file, runs the drop-in module
  ! fpm_partition_2d (library libffpm, on the GPU) on it and writes the four
  ! terms, the F_VIF density and the vorticity to a second stream file. The
  ! Python side (tests/test_gpu_fortran.py) writes the frames and checks the
  ! outputs against the oracle; this program only exercises the Fortran
  ! binding: argument association, the logical solid mask and slip_body.
  !
  ! Input stream: int32 nx, ny, slip (0/1), mask(nx*ny) (nonzero = solid);
  ! float64 phi(nx*ny), u(nx*ny), v(nx*ny), xf(nx+1), yf(ny+1), g(2), kin(6),
  ! ring((2nx+2ny)*2*2), dt, nu. Output stream: float64 terms(4),
  ! vif_density(nx*ny), omega(nx*ny). Arrays in Fortran order.
  use,intrinsic :: iso_fortran_env,only:real64,int32,output_unit,error_unit
  use fpm_partition_2d
  implicit none
  character(len=4096)::path_in,path_out
  integer(int32)::nx,ny,slip
  integer(int32),allocatable::mask(:,:)
  real(real64),allocatable::phi(:,:),u(:,:),v(:,:),xf(:),yf(:),ring(:,:,:)
  real(real64),allocatable::vif(:,:),omega(:,:)
  real(real64)::g(2),kin(6),dt,nu,terms(4)
  integer::unit_in,unit_out

  if(command_argument_count()/=2)then
    write(error_unit,'(a)')'usage: partition_frame_driver <frame.bin> <terms.bin>'
    error stop 2
  endif
  call get_command_argument(1,path_in)
  call get_command_argument(2,path_out)

  open(newunit=unit_in,file=trim(path_in),access='stream',form='unformatted',status='old',action='read')
  read(unit_in)nx,ny,slip
  allocate(mask(nx,ny),phi(nx,ny),u(nx,ny),v(nx,ny),xf(0:nx),yf(0:ny),ring(2*nx+2*ny,2,2))
  allocate(vif(nx,ny),omega(nx,ny))
  read(unit_in)mask,phi,u,v,xf,yf,g,kin,ring,dt,nu
  close(unit_in)

  if(slip/=0)then
    call fpm_partition_terms_2d(phi,g,u,v,mask/=0,xf,yf,kin,ring,dt,nu,nx,ny,terms,vif,omega, &
         slip_body=.true.)
  else
    call fpm_partition_terms_2d(phi,g,u,v,mask/=0,xf,yf,kin,ring,dt,nu,nx,ny,terms,vif,omega)
  endif

  open(newunit=unit_out,file=trim(path_out),access='stream',form='unformatted',status='replace', &
       action='write')
  write(unit_out)terms,vif,omega
  close(unit_out)
  write(output_unit,'(a,i0,a,i0,a,4es14.6)')'frame ',nx,'x',ny,' terms',terms
end program partition_frame_driver
