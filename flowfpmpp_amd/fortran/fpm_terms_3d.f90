module fpm_terms_3d
  ! Drop-in replacement of the reference module of the same name
  ! (src/fpm_terms_3d.f90:1-6): no-body FPM densities on one MI355X.
  use,intrinsic :: iso_fortran_env,only:real64
  use,intrinsic :: iso_c_binding
  use fpm_ffpm_status,only:stop_on
  implicit none
  private
  public :: compute_fpm_terms_uniform_3d
  public :: compute_fpm_terms_nonuniform_3d

  interface
    integer(c_int) function c_uniform(phi,u,v,w,ox,oy,oz,ax,ay,az,fluid,gx,gy,gz,vortex,advection,totals, &
         nx,ny,nz,dx,dy,dz) bind(C,name='ffpm_compute_fpm_terms_uniform_3d')
      import::c_int,c_double
      real(c_double)::phi(*),u(*),v(*),w(*),ox(*),oy(*),oz(*),ax(*),ay(*),az(*)
      real(c_double)::gx(*),gy(*),gz(*),vortex(*),advection(*),totals(*)
      integer(c_int)::fluid(*)
      integer(c_int),value::nx,ny,nz
      real(c_double),value::dx,dy,dz
    end function
    integer(c_int) function c_nonuniform(phi,u,v,w,ox,oy,oz,ax,ay,az,fluid,xc,yc,zc,dx,dy,dz, &
         xlo,xhi,ylo,yhi,zlo,zhi,bounds,gx,gy,gz,vortex,advection,totals,nx,ny,nz) &
         bind(C,name='ffpm_compute_fpm_terms_nonuniform_3d')
      import::c_int,c_double
      real(c_double)::phi(*),u(*),v(*),w(*),ox(*),oy(*),oz(*),ax(*),ay(*),az(*)
      real(c_double)::xc(*),yc(*),zc(*),dx(*),dy(*),dz(*),xlo(*),xhi(*),ylo(*),yhi(*),zlo(*),zhi(*)
      real(c_double)::bounds(*),gx(*),gy(*),gz(*),vortex(*),advection(*),totals(*)
      integer(c_int)::fluid(*)
      integer(c_int),value::nx,ny,nz
    end function
  end interface

contains
  ! src/fpm_terms_3d.f90:8-37
  subroutine compute_fpm_terms_uniform_3d(phi,u,v,w,ox,oy,oz,ax,ay,az,fluid,gx,gy,gz,vortex,advection,totals, &
       nx,ny,nz,dx,dy,dz)
    integer,intent(in)::nx,ny,nz,fluid(nx,ny,nz)
    real(real64),intent(in)::phi(0:nx+1,0:ny+1,0:nz+1),u(nx,ny,nz),v(nx,ny,nz),w(nx,ny,nz)
    real(real64),intent(in)::ox(nx,ny,nz),oy(nx,ny,nz),oz(nx,ny,nz),ax(nx,ny,nz),ay(nx,ny,nz),az(nx,ny,nz)
    real(real64),intent(out)::gx(nx,ny,nz),gy(nx,ny,nz),gz(nx,ny,nz),vortex(nx,ny,nz),advection(nx,ny,nz),totals(2)
    real(real64),intent(in)::dx,dy,dz
    call stop_on(c_uniform(phi,u,v,w,ox,oy,oz,ax,ay,az,fluid,gx,gy,gz,vortex,advection,totals, &
         nx,ny,nz,dx,dy,dz))
  end subroutine compute_fpm_terms_uniform_3d

  ! src/fpm_terms_3d.f90:39-87
  subroutine compute_fpm_terms_nonuniform_3d(phi,u,v,w,ox,oy,oz,ax,ay,az,fluid, &
       xc,yc,zc,dx,dy,dz,xlo,xhi,ylo,yhi,zlo,zhi,bounds, &
       gx,gy,gz,vortex,advection,totals,nx,ny,nz)
    integer,intent(in)::nx,ny,nz,fluid(nx,ny,nz)
    real(real64),intent(in)::phi(nx,ny,nz),u(nx,ny,nz),v(nx,ny,nz),w(nx,ny,nz)
    real(real64),intent(in)::ox(nx,ny,nz),oy(nx,ny,nz),oz(nx,ny,nz)
    real(real64),intent(in)::ax(nx,ny,nz),ay(nx,ny,nz),az(nx,ny,nz)
    real(real64),intent(in)::xc(nx),yc(ny),zc(nz),dx(nx),dy(ny),dz(nz)
    real(real64),intent(in)::xlo(ny,nz),xhi(ny,nz),ylo(nx,nz),yhi(nx,nz)
    real(real64),intent(in)::zlo(nx,ny),zhi(nx,ny),bounds(6)
    real(real64),intent(out)::gx(nx,ny,nz),gy(nx,ny,nz),gz(nx,ny,nz)
    real(real64),intent(out)::vortex(nx,ny,nz),advection(nx,ny,nz),totals(2)
    call stop_on(c_nonuniform(phi,u,v,w,ox,oy,oz,ax,ay,az,fluid,xc,yc,zc,dx,dy,dz, &
         xlo,xhi,ylo,yhi,zlo,zhi,bounds,gx,gy,gz,vortex,advection,totals,nx,ny,nz))
  end subroutine compute_fpm_terms_nonuniform_3d
end module fpm_terms_3d
