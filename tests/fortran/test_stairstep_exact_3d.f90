program test_stairstep_exact_3d
  ! Drop-in check of fpm_stairstep_mpi_3d on one GPU: the case of the
  ! reference's tests/test_stairstep_exact_3d.f90 (24x20x10 stretched grid,
  ! four blocked bodies, one touching the outer boundary). With body and outer
  ! face gradients g and f = 0, phi = g.x minus its fluid mean is the exact
  ! discrete solution; gradients equal g and the densities have closed forms.
  ! Jacobi, PCG and the library's multigrid-preconditioned CG ('fcg-mg'); the
  ! reference's pass criteria (exact_3d.f90:46-50). The x influence field then
  ! runs 600 fixed Jacobi sweeps, PCG and 'fcg-mg' to 1e-11.
  use,intrinsic :: iso_fortran_env,only:real64,output_unit,error_unit
  use,intrinsic :: ieee_arithmetic,only:ieee_is_finite
  use fpm_stairstep_mpi_3d
  implicit none
  integer,parameter::nx=24,ny=20,nz=10,none=-1,comm=0
  real(real64),parameter::g(3)=[0.3_real64,-0.7_real64,0.5_real64]
  real(real64)::xc(0:nx+1),yc(0:ny+1),zc(nz),dx(nx),dy(ny),dz(nz),fx(0:nx),fy(0:ny),fz(0:nz)
  real(real64)::phi(0:nx+1,0:ny+1,nz),rhs(nx,ny,nz)
  real(real64)::xlo(ny,nz),xhi(ny,nz),ylo(nx,nz),yhi(nx,nz),zlo(nx,ny),zhi(nx,ny)
  real(real64)::u(nx,ny,nz),v(nx,ny,nz),w(nx,ny,nz),ox(nx,ny,nz),oy(nx,ny,nz),oz(nx,ny,nz)
  real(real64)::ax(nx,ny,nz),ay(nx,ny,nz),az(nx,ny,nz)
  real(real64)::gx(nx,ny,nz),gy(nx,ny,nz),gz(nx,ny,nz),vortex(nx,ny,nz),advection(nx,ny,nz)
  real(real64)::body(3),bounds(6),totals(2),expect(2),corr,imb,res,chk,tol,err(4),mean,vol,fluidvol
  real(real64)::cx,cy,cz,ev,p
  integer::solid(0:nx+1,0:ny+1,nz),i,j,k,sweeps,m
  logical::ok,all_ok
  character(8),parameter::methods(3)=[character(8)::'jacobi','pcg','fcg-mg']

  call axis(nx,0.0_real64,1.2_real64,0.21_real64,fx,xc(1:nx),dx)
  call axis(ny,0.0_real64,1.0_real64,-0.17_real64,fy,yc(1:ny),dy)
  call axis(nz,0.0_real64,0.5_real64,0.13_real64,fz,zc,dz)
  xc(0)=xc(1)-dx(1);xc(nx+1)=xc(nx)+dx(nx)
  yc(0)=yc(1)-dy(1);yc(ny+1)=yc(ny)+dy(ny)
  solid=0
  do k=1,nz;do j=1,ny;do i=1,nx
    if(blocked(i,j,k))solid(i,j,k)=1
  enddo;enddo;enddo
  call exchange_solid_halo_3d(solid,nx,ny,nz,none,none,none,none,comm)
  do k=1,nz;do j=1,ny;do i=1,nx
    u(i,j,k)=1+sin(3*xc(i))*cos(2*yc(j));v(i,j,k)=0.5_real64*cos(xc(i)+zc(k));w(i,j,k)=0.2_real64*yc(j)
    ox(i,j,k)=cos(4*zc(k));oy(i,j,k)=-0.3_real64+xc(i)*yc(j);oz(i,j,k)=sin(5*yc(j))
    ax(i,j,k)=0.1_real64*xc(i);ay(i,j,k)=-0.2_real64;az(i,j,k)=zc(k)*yc(j)
  enddo;enddo;enddo
  bounds=[-1.0_real64,2.0_real64,-1.0_real64,2.0_real64,-1.0_real64,1.0_real64]
  all_ok=.true.

  do m=1,3
    tol=merge(1e-12_real64,1e-10_real64,m==1)
    body=g;xlo=g(1);xhi=g(1);ylo=g(2);yhi=g(2);zlo=g(3);zhi=g(3);rhs=0;phi=0
    call enforce_stairstep_compatibility_3d(rhs,solid,body,xlo,xhi,ylo,yhi,zlo,zhi,dx,dy,dz, &
         nx,ny,nz,none,none,none,none,comm,corr,imb)
    call solve_stairstep_poisson_3d(phi,rhs,solid,body,xlo,xhi,ylo,yhi,zlo,zhi,xc,yc,zc,dx,dy,dz, &
         nx,ny,nz,none,none,none,none,comm,tol,400000,50,sweeps,res,use_openacc=.true.,method=trim(methods(m)))
    chk=stairstep_residual_3d(phi,rhs,solid,body,xlo,xhi,ylo,yhi,zlo,zhi,xc,yc,zc,dx,dy,dz, &
         nx,ny,nz,none,none,none,none,comm)
    call compute_stairstep_fpm_terms_3d(phi,u,v,w,ox,oy,oz,ax,ay,az,solid,body,xc,yc,zc,dx,dy,dz, &
         xlo,xhi,ylo,yhi,zlo,zhi,bounds,gx,gy,gz,vortex,advection,totals,nx,ny,nz,none,none,none,none,comm)
    mean=0;fluidvol=0
    do k=1,nz;do j=1,ny;do i=1,nx
      if(solid(i,j,k)/=0)cycle
      vol=dx(i)*dy(j)*dz(k)
      mean=mean+vol*(g(1)*xc(i)+g(2)*yc(j)+g(3)*zc(k));fluidvol=fluidvol+vol
    enddo;enddo;enddo
    mean=mean/fluidvol
    err=0;expect=0
    do k=1,nz;do j=1,ny;do i=1,nx
      if(solid(i,j,k)/=0)then
        err(4)=max(err(4),abs(phi(i,j,k)),abs(vortex(i,j,k)),abs(gx(i,j,k)))
        cycle
      endif
      p=g(1)*xc(i)+g(2)*yc(j)+g(3)*zc(k)-mean
      err(1)=max(err(1),abs(phi(i,j,k)-p))
      err(2)=max(err(2),abs(gx(i,j,k)-g(1)),abs(gy(i,j,k)-g(2)),abs(gz(i,j,k)-g(3)))
      cx=w(i,j,k)*oy(i,j,k)-v(i,j,k)*oz(i,j,k)
      cy=u(i,j,k)*oz(i,j,k)-w(i,j,k)*ox(i,j,k)
      cz=v(i,j,k)*ox(i,j,k)-u(i,j,k)*oy(i,j,k)
      ev=-(cx*g(1)+cy*g(2)+cz*g(3))
      err(3)=max(err(3),abs(vortex(i,j,k)-ev))
      vol=dx(i)*dy(j)*dz(k)
      expect=expect+vol*[ev,-(ax(i,j,k)*g(1)+ay(i,j,k)*g(2)+az(i,j,k)*g(3))]
    enddo;enddo;enddo
    ok=abs(corr)<1e-13.and.abs(imb)<1e-13.and.res<=2*tol.and.chk<1e-9
    ok=ok.and.err(1)<1e-9.and.err(2)<1e-9.and.err(3)<1e-9.and.err(4)==0
    ok=ok.and.maxval(abs(totals-expect))<1e-10
    ok=ok.and.all(ieee_is_finite([res,chk,err,totals]))
    write(output_unit,'(a,a,a,i0,a,2es12.4,a,4es12.4)')'exact ',trim(methods(m)),': iterations ',sweeps, &
         ', residual/check ',res,chk,', errors ',err
    all_ok=all_ok.and.ok

    ! influence field of direction x
    body=[1.0_real64,0.0_real64,0.0_real64];xlo=0;xhi=0;ylo=0;yhi=0;zlo=0;zhi=0;rhs=0;phi=0
    call enforce_stairstep_compatibility_3d(rhs,solid,body,xlo,xhi,ylo,yhi,zlo,zhi,dx,dy,dz, &
         nx,ny,nz,none,none,none,none,comm,corr,imb)
    if(m==1)then
      call solve_stairstep_poisson_3d(phi,rhs,solid,body,xlo,xhi,ylo,yhi,zlo,zhi,xc,yc,zc,dx,dy,dz, &
           nx,ny,nz,none,none,none,none,comm,-1.0_real64,600,50,sweeps,res,method='jacobi')
      ok=sweeps==600
    else
      call solve_stairstep_poisson_3d(phi,rhs,solid,body,xlo,xhi,ylo,yhi,zlo,zhi,xc,yc,zc,dx,dy,dz, &
           nx,ny,nz,none,none,none,none,comm,1e-11_real64,100000,50,sweeps,res,method=trim(methods(m)))
      ok=res<=2e-11_real64
    endif
    chk=stairstep_residual_3d(phi,rhs,solid,body,xlo,xhi,ylo,yhi,zlo,zhi,xc,yc,zc,dx,dy,dz, &
         nx,ny,nz,none,none,none,none,comm)
    ok=ok.and.abs(imb)<1e-13.and.abs(chk-res)<=1e-8*res+1e-12.and.all(ieee_is_finite(phi))
    write(output_unit,'(a,a,a,i0,a,es12.4,a,es12.4)')'influence ',trim(methods(m)),': iterations ',sweeps, &
         ', residual ',res,', correction ',corr
    all_ok=all_ok.and.ok
  enddo
  if(.not.all_ok)then
    write(error_unit,'(a)')'FAIL: stair-step exactness test (drop-in module, GPU)'
    error stop 1
  endif
  write(output_unit,'(a)')'PASS: stair-step exactness test (drop-in module, GPU)'
contains
  logical function blocked(i,j,k)
    integer,intent(in)::i,j,k
    blocked=(i>=6.and.i<=10.and.j>=5.and.j<=12.and.k>=3.and.k<=7) &
         .or.(i>=15.and.i<=19.and.j>=12.and.j<=16.and.k>=2.and.k<=9) &
         .or.(i>=15.and.i<=17.and.j>=8.and.j<=11.and.k>=2.and.k<=9) &
         .or.(i>=22.and.j<=3)
  end function blocked

  ! Axis on [lo,hi] with widths proportional to 1 + a*sin(2 pi (m-1/2)/n)
  ! (the reference's test-grid axis, tests/fpm_test_grid.f90:11-27).
  subroutine axis(n,lo,hi,a,face,center,width)
    integer,intent(in)::n
    real(real64),intent(in)::lo,hi,a
    real(real64),intent(out)::face(0:n),center(n),width(n)
    real(real64)::wt(n),total
    integer::m
    do m=1,n
      wt(m)=1+a*sin(2*acos(-1.0_real64)*(m-0.5_real64)/n)
    enddo
    total=sum(wt)
    face(0)=lo
    do m=1,n
      width(m)=(hi-lo)*wt(m)/total
      face(m)=face(m-1)+width(m)
      center(m)=0.5_real64*(face(m-1)+face(m))
    enddo
    face(n)=hi
  end subroutine axis
end program test_stairstep_exact_3d
