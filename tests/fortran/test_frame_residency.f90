program test_frame_residency
  ! The reference's frame sequence -- compatibility correction, influence
  ! solve, force densities (tests/benchmark_stairstep_3d.f90:151-165) -- run
  ! twice through the drop-in module fpm_stairstep_mpi_3d: first on plain host
  ! arrays (every call stages its arrays), then with every array bound to a
  ! frame of fpm_ffpm_frame (the OpenACC data region's role). Pass criteria:
  ! the two runs agree bit for bit (phi, corrected faces, the five density
  ! fields, the totals); inside the frame no call moves an array (the
  ! library's transfer counters do not change); binding uploads each array
  ! once and ending the frame downloads each copy_out array once.
  use,intrinsic :: iso_fortran_env,only:real64,int64,output_unit
  use fpm_stairstep_mpi_3d
  use fpm_ffpm_frame
  implicit none
  integer,parameter::nx=40,ny=32,nz=24,null=-1,comm=0
  real(real64),parameter::pi=3.141592653589793_real64
  integer::solid(0:nx+1,0:ny+1,nz),i,j,k,sweeps(2),nbound,ncopyout
  real(real64)::phi(0:nx+1,0:ny+1,nz,2),rhs(nx,ny,nz)
  real(real64)::xlo(ny,nz,2),xhi(ny,nz,2),ylo(nx,nz,2),yhi(nx,nz,2),zlo(nx,ny,2),zhi(nx,ny,2)
  real(real64)::xc(0:nx+1),yc(0:ny+1),zc(nz),dx(nx),dy(ny),dz(nz),bg(3),bounds(6)
  real(real64),dimension(nx,ny,nz)::u,v,w,ox,oy,oz,ax,ay,az
  real(real64)::gx(nx,ny,nz,2),gy(nx,ny,nz,2),gz(nx,ny,nz,2),vo(nx,ny,nz,2),ad(nx,ny,nz,2)
  real(real64)::totals(2,2),corr(2),imb(2),res(2)
  integer(int64)::c0(4),c1(4),c2(4),c3(4)
  type(ffpm_frame)::frame
  logical::ok

  call axis(nx,-4.0_real64,8.0_real64,dx,xc(1:nx))
  call axis(ny,-3.0_real64,6.0_real64,dy,yc(1:ny))
  call axis(nz,-2.0_real64,4.0_real64,dz,zc)
  xc(0)=xc(1)-dx(1);xc(nx+1)=xc(nx)+dx(nx)
  yc(0)=yc(1)-dy(1);yc(ny+1)=yc(ny)+dy(ny)
  solid=0
  do k=1,nz;do j=1,ny;do i=1,nx
    if(xc(i)**2+yc(j)**2+zc(k)**2<1.1_real64**2)solid(i,j,k)=1
    if(xc(i)>3.2_real64.and.abs(yc(j))<0.9_real64.and.abs(zc(k))<0.6_real64)solid(i,j,k)=1
    u(i,j,k)=sin(xc(i))*cos(yc(j));v(i,j,k)=-cos(xc(i))*sin(yc(j));w(i,j,k)=0.3_real64*zc(k)
    ox(i,j,k)=cos(zc(k));oy(i,j,k)=sin(xc(i)+yc(j));oz(i,j,k)=0.5_real64
    ax(i,j,k)=xc(i)*yc(j);ay(i,j,k)=zc(k);az(i,j,k)=-0.2_real64
  enddo;enddo;enddo
  rhs=0;bg=[1.0_real64,0.0_real64,0.0_real64]
  bounds=[-4.0_real64,4.0_real64,-3.0_real64,3.0_real64,-2.0_real64,2.0_real64]*1.000001_real64
  phi=0;xlo=0;xhi=0;ylo=0;yhi=0;zlo=0;zhi=0

  ! 1. plain host arrays; use_openacc=.false. must print the backend notice
  call ffpm_transfer_counts(c0)
  call run(1,.false.)
  call ffpm_transfer_counts(c1)

  ! 2. the same sequence inside a frame
  call ffpm_frame_begin(frame)
  call ffpm_frame_bind(frame,phi(:,:,:,2),copy_out=.true.)
  call ffpm_frame_bind(frame,rhs)
  call ffpm_frame_bind(frame,solid)
  call ffpm_frame_bind(frame,xlo(:,:,2),copy_out=.true.)
  call ffpm_frame_bind(frame,xhi(:,:,2),copy_out=.true.)
  call ffpm_frame_bind(frame,ylo(:,:,2),copy_out=.true.)
  call ffpm_frame_bind(frame,yhi(:,:,2),copy_out=.true.)
  call ffpm_frame_bind(frame,zlo(:,:,2),copy_out=.true.)
  call ffpm_frame_bind(frame,zhi(:,:,2),copy_out=.true.)
  call ffpm_frame_bind(frame,xc)
  call ffpm_frame_bind(frame,yc)
  call ffpm_frame_bind(frame,zc)
  call ffpm_frame_bind(frame,dx)
  call ffpm_frame_bind(frame,dy)
  call ffpm_frame_bind(frame,dz)
  call ffpm_frame_bind(frame,u)
  call ffpm_frame_bind(frame,v)
  call ffpm_frame_bind(frame,w)
  call ffpm_frame_bind(frame,ox)
  call ffpm_frame_bind(frame,oy)
  call ffpm_frame_bind(frame,oz)
  call ffpm_frame_bind(frame,ax)
  call ffpm_frame_bind(frame,ay)
  call ffpm_frame_bind(frame,az)
  call ffpm_frame_bind(frame,gx(:,:,:,2),copy_out=.true.)
  call ffpm_frame_bind(frame,gy(:,:,:,2),copy_out=.true.)
  call ffpm_frame_bind(frame,gz(:,:,:,2),copy_out=.true.)
  call ffpm_frame_bind(frame,vo(:,:,:,2),copy_out=.true.)
  call ffpm_frame_bind(frame,ad(:,:,:,2),copy_out=.true.)
  nbound=29;ncopyout=12
  call ffpm_transfer_counts(c2)
  call run(2,.true.)
  call ffpm_transfer_counts(c3)
  ok=all(c3==c2)
  write(output_unit,'(a,4i12)')'transfers inside the frame (uploads, bytes, downloads, bytes):',c3-c2
  call ffpm_frame_end(frame)
  call ffpm_transfer_counts(c3)
  write(output_unit,'(a,4i12)')'transfers of the unbound run:                                 ',c1-c0
  write(output_unit,'(a,2i6)')'frame uploads at bind, downloads at end:',c2(1)-c1(1),c3(3)-c2(3)
  ok=ok.and.c2(1)-c1(1)==nbound.and.c3(3)-c2(3)==ncopyout.and.c2(3)==c1(3).and.c3(1)==c2(1)
  ok=ok.and.c1(1)-c0(1)>=nbound
  write(output_unit,'(a,2i6,2es12.4)')'PCG iterations and residuals:',sweeps,res
  ok=ok.and.sweeps(1)==sweeps(2).and.res(1)==res(2).and.res(1)<=1e-10_real64
  ok=ok.and.corr(1)==corr(2).and.imb(1)==imb(2)
  ok=ok.and.all(phi(:,:,:,1)==phi(:,:,:,2))
  ok=ok.and.all(xlo(:,:,1)==xlo(:,:,2)).and.all(xhi(:,:,1)==xhi(:,:,2))
  ok=ok.and.all(ylo(:,:,1)==ylo(:,:,2)).and.all(yhi(:,:,1)==yhi(:,:,2))
  ok=ok.and.all(zlo(:,:,1)==zlo(:,:,2)).and.all(zhi(:,:,1)==zhi(:,:,2))
  ok=ok.and.all(gx(:,:,:,1)==gx(:,:,:,2)).and.all(gy(:,:,:,1)==gy(:,:,:,2))
  ok=ok.and.all(gz(:,:,:,1)==gz(:,:,:,2)).and.all(vo(:,:,:,1)==vo(:,:,:,2))
  ok=ok.and.all(ad(:,:,:,1)==ad(:,:,:,2)).and.all(totals(:,1)==totals(:,2))
  ok=ok.and.maxval(abs(phi(:,:,:,2)))>0.and.maxval(abs(vo(:,:,:,2)))>0
  if(ok)then
    write(output_unit,'(a)')'PASS'
  else
    write(output_unit,'(a)')'FAIL'
    error stop 1
  endif

contains
  ! cell widths proportional to 1 + 0.4 cos(2 pi s) over [x0, x0 + length],
  ! centres at the face midpoints
  subroutine axis(n,x0,length,d,c)
    integer,intent(in)::n
    real(real64),intent(in)::x0,length
    real(real64),intent(out)::d(n),c(n)
    real(real64)::f
    integer::m
    do m=1,n
      d(m)=1+0.4_real64*cos(2*pi*(m-0.5_real64)/n)
    enddo
    d=d*length/sum(d)
    f=x0
    do m=1,n
      c(m)=f+0.5_real64*d(m)
      f=f+d(m)
    enddo
  end subroutine axis

  subroutine run(s,gpu)
    integer,intent(in)::s
    logical,intent(in)::gpu
    call enforce_stairstep_compatibility_3d(rhs,solid,bg,xlo(:,:,s),xhi(:,:,s),ylo(:,:,s), &
         yhi(:,:,s),zlo(:,:,s),zhi(:,:,s),dx,dy,dz,nx,ny,nz,null,null,null,null,comm,corr(s),imb(s))
    call solve_stairstep_poisson_3d(phi(:,:,:,s),rhs,solid,bg,xlo(:,:,s),xhi(:,:,s),ylo(:,:,s), &
         yhi(:,:,s),zlo(:,:,s),zhi(:,:,s),xc,yc,zc,dx,dy,dz,nx,ny,nz,null,null,null,null,comm, &
         1e-10_real64,100000,50,sweeps(s),res(s),use_openacc=gpu)
    call compute_stairstep_fpm_terms_3d(phi(:,:,:,s),u,v,w,ox,oy,oz,ax,ay,az,solid,bg,xc,yc,zc, &
         dx,dy,dz,xlo(:,:,s),xhi(:,:,s),ylo(:,:,s),yhi(:,:,s),zlo(:,:,s),zhi(:,:,s),bounds, &
         gx(:,:,:,s),gy(:,:,:,s),gz(:,:,:,s),vo(:,:,:,s),ad(:,:,:,s),totals(:,s),nx,ny,nz, &
         null,null,null,null,comm,use_openacc=gpu)
  end subroutine run
end program test_frame_residency
