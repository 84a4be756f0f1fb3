module fpm_laplace_3d
  ! Drop-in replacement of the reference module of the same name
  ! (src/fpm_laplace_3d.f90:1-8): the no-body pure-Neumann compatibility
  ! corrections and Jacobi solvers, run on one MI355X through libffpm's C ABI.
  use,intrinsic :: iso_fortran_env,only:real64
  use,intrinsic :: iso_c_binding
  use fpm_ffpm_status,only:stop_on
  implicit none
  private
  public :: enforce_neumann_compatibility_3d,solve_neumann_poisson_3d
  public :: enforce_neumann_compatibility_nonuniform_3d
  public :: solve_neumann_poisson_nonuniform_3d

  interface
    integer(c_int) function c_compat(rhs,xlo,xhi,ylo,yhi,zlo,zhi,nx,ny,nz,dx,dy,dz,correction,imbalance) &
         bind(C,name='ffpm_enforce_neumann_compatibility_3d')
      import::c_int,c_double
      real(c_double)::rhs(*),xlo(*),xhi(*),ylo(*),yhi(*),zlo(*),zhi(*),correction,imbalance
      integer(c_int),value::nx,ny,nz
      real(c_double),value::dx,dy,dz
    end function
    integer(c_int) function c_solve(phi,rhs,xlo,xhi,ylo,yhi,zlo,zhi,nx,ny,nz,dx,dy,dz,tolerance, &
         max_iterations,check_interval,iterations,final_residual) bind(C,name='ffpm_solve_neumann_poisson_3d')
      import::c_int,c_double
      real(c_double)::phi(*),rhs(*),xlo(*),xhi(*),ylo(*),yhi(*),zlo(*),zhi(*),final_residual
      integer(c_int),value::nx,ny,nz,max_iterations,check_interval
      real(c_double),value::dx,dy,dz,tolerance
      integer(c_int)::iterations
    end function
    integer(c_int) function c_compat_nu(rhs,xlo,xhi,ylo,yhi,zlo,zhi,dx,dy,dz,nx,ny,nz,correction,imbalance) &
         bind(C,name='ffpm_enforce_neumann_compatibility_nonuniform_3d')
      import::c_int,c_double
      real(c_double)::rhs(*),xlo(*),xhi(*),ylo(*),yhi(*),zlo(*),zhi(*),dx(*),dy(*),dz(*)
      real(c_double)::correction,imbalance
      integer(c_int),value::nx,ny,nz
    end function
    integer(c_int) function c_solve_nu(phi,rhs,xlo,xhi,ylo,yhi,zlo,zhi,xc,yc,zc,dx,dy,dz,nx,ny,nz, &
         tolerance,max_iterations,check_interval,iterations,final_residual) &
         bind(C,name='ffpm_solve_neumann_poisson_nonuniform_3d')
      import::c_int,c_double
      real(c_double)::phi(*),rhs(*),xlo(*),xhi(*),ylo(*),yhi(*),zlo(*),zhi(*)
      real(c_double)::xc(*),yc(*),zc(*),dx(*),dy(*),dz(*),final_residual
      integer(c_int),value::nx,ny,nz,max_iterations,check_interval
      real(c_double),value::tolerance
      integer(c_int)::iterations
    end function
  end interface

contains
  ! src/fpm_laplace_3d.f90:9-53
  subroutine enforce_neumann_compatibility_3d(rhs,xlo,xhi,ylo,yhi,zlo,zhi,nx,ny,nz,dx,dy,dz,correction,imbalance)
    integer,intent(in)::nx,ny,nz
    real(real64),intent(in)::rhs(nx,ny,nz),dx,dy,dz
    real(real64),intent(inout)::xlo(ny,nz),xhi(ny,nz),ylo(nx,nz),yhi(nx,nz),zlo(nx,ny),zhi(nx,ny)
    real(real64),intent(out)::correction,imbalance
    call stop_on(c_compat(rhs,xlo,xhi,ylo,yhi,zlo,zhi,nx,ny,nz,dx,dy,dz,correction,imbalance))
  end subroutine enforce_neumann_compatibility_3d

  ! src/fpm_laplace_3d.f90:55-107
  subroutine solve_neumann_poisson_3d(phi,rhs,xlo,xhi,ylo,yhi,zlo,zhi,nx,ny,nz,dx,dy,dz,tolerance, &
       max_iterations,check_interval,iterations,final_residual)
    integer,intent(in)::nx,ny,nz,max_iterations,check_interval
    real(real64),intent(inout)::phi(0:nx+1,0:ny+1,0:nz+1)
    real(real64),intent(in)::rhs(nx,ny,nz),xlo(ny,nz),xhi(ny,nz),ylo(nx,nz),yhi(nx,nz),zlo(nx,ny),zhi(nx,ny)
    real(real64),intent(in)::dx,dy,dz,tolerance
    integer,intent(out)::iterations
    real(real64),intent(out)::final_residual
    call stop_on(c_solve(phi,rhs,xlo,xhi,ylo,yhi,zlo,zhi,nx,ny,nz,dx,dy,dz,tolerance, &
         max_iterations,check_interval,iterations,final_residual))
  end subroutine solve_neumann_poisson_3d

  ! src/fpm_laplace_3d.f90:150-217
  subroutine enforce_neumann_compatibility_nonuniform_3d(rhs,xlo,xhi,ylo,yhi,zlo,zhi, &
       dx,dy,dz,nx,ny,nz,correction,imbalance)
    integer,intent(in)::nx,ny,nz
    real(real64),intent(in)::rhs(nx,ny,nz),dx(nx),dy(ny),dz(nz)
    real(real64),intent(inout)::xlo(ny,nz),xhi(ny,nz),ylo(nx,nz),yhi(nx,nz),zlo(nx,ny),zhi(nx,ny)
    real(real64),intent(out)::correction,imbalance
    call stop_on(c_compat_nu(rhs,xlo,xhi,ylo,yhi,zlo,zhi,dx,dy,dz,nx,ny,nz,correction,imbalance))
  end subroutine enforce_neumann_compatibility_nonuniform_3d

  ! src/fpm_laplace_3d.f90:219-293
  subroutine solve_neumann_poisson_nonuniform_3d(phi,rhs,xlo,xhi,ylo,yhi,zlo,zhi, &
       xc,yc,zc,dx,dy,dz,nx,ny,nz,tolerance,max_iterations,check_interval,iterations,final_residual)
    integer,intent(in)::nx,ny,nz,max_iterations,check_interval
    real(real64),intent(inout)::phi(nx,ny,nz)
    real(real64),intent(in)::rhs(nx,ny,nz),xlo(ny,nz),xhi(ny,nz),ylo(nx,nz),yhi(nx,nz)
    real(real64),intent(in)::zlo(nx,ny),zhi(nx,ny),xc(nx),yc(ny),zc(nz),dx(nx),dy(ny),dz(nz),tolerance
    integer,intent(out)::iterations
    real(real64),intent(out)::final_residual
    call stop_on(c_solve_nu(phi,rhs,xlo,xhi,ylo,yhi,zlo,zhi,xc,yc,zc,dx,dy,dz,nx,ny,nz, &
         tolerance,max_iterations,check_interval,iterations,final_residual))
  end subroutine solve_neumann_poisson_nonuniform_3d
end module fpm_laplace_3d
