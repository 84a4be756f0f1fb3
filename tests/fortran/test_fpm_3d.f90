program test_fpm_3d
  ! Drop-in check of fpm_laplace_3d and fpm_terms_3d on one GPU with the
  ! manufactured quadratic of the reference's tests/test_fpm_3d.f90: uniform
  ! 12x10x8 grid, rhs = 1, zero face data, so the compatibility correction is
  ! 1/6 and phi = |x - 1/2|^2/6 (minus its mean) is exact; the reference's
  ! pass criteria (test_fpm_3d.f90:47-48).
  use,intrinsic :: iso_fortran_env,only:real64,output_unit,error_unit
  use fpm_laplace_3d
  use fpm_terms_3d
  implicit none
  integer,parameter::nx=12,ny=10,nz=8
  real(real64)::phi(0:nx+1,0:ny+1,0:nz+1),rhs(nx,ny,nz),exact(nx,ny,nz)
  real(real64)::u(nx,ny,nz),v(nx,ny,nz),w(nx,ny,nz),ox(nx,ny,nz),oy(nx,ny,nz),oz(nx,ny,nz)
  real(real64)::ax(nx,ny,nz),ay(nx,ny,nz),az(nx,ny,nz)
  real(real64)::gx(nx,ny,nz),gy(nx,ny,nz),gz(nx,ny,nz),vf(nx,ny,nz),af(nx,ny,nz)
  real(real64)::xlo(ny,nz),xhi(ny,nz),ylo(nx,nz),yhi(nx,nz),zlo(nx,ny),zhi(nx,ny)
  real(real64)::dx,dy,dz,x,y,z,corr,imb,res,totals(2),e_sol,e_grad,e_field,ev,ea
  integer::fluid(nx,ny,nz),i,j,k,iters
  logical::ok
  dx=1.0_real64/nx;dy=1.0_real64/ny;dz=1.0_real64/nz
  phi=0;rhs=1;xlo=0;xhi=0;ylo=0;yhi=0;zlo=0;zhi=0
  call enforce_neumann_compatibility_3d(rhs,xlo,xhi,ylo,yhi,zlo,zhi,nx,ny,nz,dx,dy,dz,corr,imb)
  do k=1,nz;do j=1,ny;do i=1,nx
    x=(i-0.5_real64)*dx;y=(j-0.5_real64)*dy;z=(k-0.5_real64)*dz
    exact(i,j,k)=((x-0.5_real64)**2+(y-0.5_real64)**2+(z-0.5_real64)**2)/6
  enddo;enddo;enddo
  exact=exact-sum(exact)/(nx*ny*nz)
  call solve_neumann_poisson_3d(phi,rhs,xlo,xhi,ylo,yhi,zlo,zhi,nx,ny,nz,dx,dy,dz,1e-10_real64,50000,25,iters,res)
  u=1;v=2;w=3;ox=0.5_real64;oy=-0.25_real64;oz=0.75_real64;ax=0.2_real64;ay=-0.1_real64;az=0.3_real64;fluid=1
  call compute_fpm_terms_uniform_3d(phi,u,v,w,ox,oy,oz,ax,ay,az,fluid,gx,gy,gz,vf,af,totals,nx,ny,nz,dx,dy,dz)
  e_sol=maxval(abs(phi(1:nx,1:ny,1:nz)-exact));e_grad=0;e_field=0
  do k=1,nz;do j=1,ny;do i=1,nx
    x=(i-0.5_real64)*dx-0.5_real64;y=(j-0.5_real64)*dy-0.5_real64;z=(k-0.5_real64)*dz-0.5_real64
    e_grad=max(e_grad,abs(gx(i,j,k)-x/3),abs(gy(i,j,k)-y/3),abs(gz(i,j,k)-z/3))
    ! omega x u = (-2.25, -0.75, 1.25) for these constant fields
    ev=-(-2.25_real64*x/3-0.75_real64*y/3+1.25_real64*z/3)
    ea=-(0.2_real64*x/3-0.1_real64*y/3+0.3_real64*z/3)
    e_field=max(e_field,abs(vf(i,j,k)-ev),abs(af(i,j,k)-ea))
  enddo;enddo;enddo
  write(output_unit,'(a,2es12.4,a,i0,a,es12.4,a,3es12.4)')'correction/imbalance ',corr,imb,', iterations ',iters, &
       ', residual ',res,', errors ',e_sol,e_grad,e_field
  ok=iters<50000.and.res<=1e-10_real64.and.abs(corr-1.0_real64/6)<1e-12.and.abs(imb)<1e-12
  ok=ok.and.e_sol<1e-9.and.e_grad<1e-9.and.e_field<1e-9.and.maxval(abs(totals))<1e-12
  if(.not.ok)then
    write(error_unit,'(a)')'FAIL: manufactured 3-D Laplace/FPM test (drop-in modules, GPU)'
    error stop 1
  endif
  write(output_unit,'(a)')'PASS: manufactured 3-D Laplace/FPM test (drop-in modules, GPU)'
end program test_fpm_3d
