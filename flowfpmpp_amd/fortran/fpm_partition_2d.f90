module fpm_partition_2d
  ! Drop-in replacement of the reference module of the same name
  ! (src/fpm_partition_2d.f90:1-35): the force partition F_KINE, F_VIF, F_VIS,
  ! F_BOUND of a 2-D stair-step body, per unit span and unit density, computed
  ! on one MI355X through the C ABI of libffpm (include/ffpm.h). The logical
  ! solid mask is passed to the library as int32 (1 = solid); the optional
  ! slip_body keeps its meaning.
  use,intrinsic :: iso_fortran_env,only:real64
  use,intrinsic :: iso_c_binding
  use fpm_ffpm_status,only:stop_on
  implicit none
  private
  public::fpm_partition_terms_2d

  interface
    integer(c_int) function c_partition(phi,g,u,v,solid,xf,yf,kin,ring,dt,nu,nx,ny,slip_body, &
         terms,vif_density,omega) bind(C,name='ffpm_partition_terms_2d')
      import::c_int,c_double
      real(c_double)::phi(*),g(*),u(*),v(*),xf(*),yf(*),kin(*),ring(*),terms(*)
      real(c_double)::vif_density(*),omega(*)
      integer(c_int)::solid(*)
      real(c_double),value::dt,nu
      integer(c_int),value::nx,ny,slip_body
    end function
  end interface

contains
  ! src/fpm_partition_2d.f90:35-154
  subroutine fpm_partition_terms_2d(phi,g,u,v,solid,xf,yf,kin,ring,dt,nu,nx,ny,terms,vif_density,omega,slip_body)
    integer,intent(in)::nx,ny
    real(real64),intent(in)::phi(nx,ny),g(2),u(nx,ny),v(nx,ny),xf(0:nx),yf(0:ny),kin(6)
    real(real64),intent(in)::ring(2*nx+2*ny,2,2),dt,nu
    logical,intent(in)::solid(nx,ny)
    real(real64),intent(out)::terms(4),vif_density(nx,ny),omega(nx,ny)
    logical,intent(in),optional::slip_body
    integer(c_int),allocatable::mask(:,:)
    integer(c_int)::slip
    slip=0
    if(present(slip_body))then;if(slip_body)slip=1;endif
    allocate(mask(nx,ny))
    mask=merge(1_c_int,0_c_int,solid)
    call stop_on(c_partition(phi,g,u,v,mask,xf,yf,kin,ring,dt,nu,int(nx,c_int),int(ny,c_int),slip, &
         terms,vif_density,omega))
  end subroutine fpm_partition_terms_2d
end module fpm_partition_2d
