module fpm_stairstep_mpi_3d
  ! Drop-in replacement of the reference module of the same name
  ! (src/fpm_stairstep_mpi_3d.f90:1-22): the same public routines with the same
  ! argument lists, running on one MI355X through the C ABI of libffpm
  ! (include/ffpm.h). One GPU holds the whole domain, so every neighbour rank
  ! must be MPI_PROC_NULL (any negative value); comm is accepted and unused.
  ! use_openacc is accepted for source compatibility: the work always runs on
  ! the GPU, and use_openacc=.false. prints a one-time notice on standard
  ! error saying so (a caller's CPU-against-GPU check would otherwise compare
  ! the GPU with itself without knowing). Errors of the reference's own checks
  ! stop with the reference's message text.
  use,intrinsic :: iso_fortran_env,only:real64,error_unit
  use,intrinsic :: iso_c_binding
  use fpm_ffpm_status,only:stop_on
  implicit none
  private
  public::exchange_solid_halo_3d,enforce_stairstep_compatibility_3d,solve_stairstep_poisson_3d
  public::stairstep_residual_3d,compute_stairstep_fpm_terms_3d

  logical,save::backend_noticed=.false.

  interface
    integer(c_int) function c_exchange_solid_halo(solid,nx,ny,nz,west,east,south,north) &
         bind(C,name='ffpm_exchange_solid_halo_3d')
      import::c_int
      integer(c_int)::solid(*)
      integer(c_int),value::nx,ny,nz,west,east,south,north
    end function
    integer(c_int) function c_compat(rhs,solid,body_grad,xlo,xhi,ylo,yhi,zlo,zhi,dx,dy,dz, &
         nx,ny,nz,west,east,south,north,correction,imbalance) &
         bind(C,name='ffpm_enforce_stairstep_compatibility_3d')
      import::c_int,c_double
      real(c_double)::rhs(*),body_grad(*),xlo(*),xhi(*),ylo(*),yhi(*),zlo(*),zhi(*),dx(*),dy(*),dz(*)
      integer(c_int)::solid(*)
      integer(c_int),value::nx,ny,nz,west,east,south,north
      real(c_double)::correction,imbalance
    end function
    integer(c_int) function c_solve(phi,rhs,solid,body_grad,xlo,xhi,ylo,yhi,zlo,zhi,xc,yc,zc, &
         dx,dy,dz,nx,ny,nz,west,east,south,north,tolerance,max_sweeps,check_interval,method, &
         sweeps,residual) bind(C,name='ffpm_solve_stairstep_poisson_3d')
      import::c_int,c_double
      real(c_double)::phi(*),rhs(*),body_grad(*),xlo(*),xhi(*),ylo(*),yhi(*),zlo(*),zhi(*)
      real(c_double)::xc(*),yc(*),zc(*),dx(*),dy(*),dz(*)
      integer(c_int)::solid(*)
      integer(c_int),value::nx,ny,nz,west,east,south,north,max_sweeps,check_interval,method
      real(c_double),value::tolerance
      integer(c_int)::sweeps
      real(c_double)::residual
    end function
    integer(c_int) function c_residual(phi,rhs,solid,body_grad,xlo,xhi,ylo,yhi,zlo,zhi,xc,yc,zc, &
         dx,dy,dz,nx,ny,nz,west,east,south,north,value) bind(C,name='ffpm_stairstep_residual_3d')
      import::c_int,c_double
      real(c_double)::phi(*),rhs(*),body_grad(*),xlo(*),xhi(*),ylo(*),yhi(*),zlo(*),zhi(*)
      real(c_double)::xc(*),yc(*),zc(*),dx(*),dy(*),dz(*)
      integer(c_int)::solid(*)
      integer(c_int),value::nx,ny,nz,west,east,south,north
      real(c_double)::value
    end function
    integer(c_int) function c_terms(phi,u,v,w,ox,oy,oz,ax,ay,az,solid,body_grad,xc,yc,zc,dx,dy,dz, &
         xlo,xhi,ylo,yhi,zlo,zhi,bounds,gx,gy,gz,vortex,advection,totals,nx,ny,nz, &
         west,east,south,north) bind(C,name='ffpm_compute_stairstep_fpm_terms_3d')
      import::c_int,c_double
      real(c_double)::phi(*),u(*),v(*),w(*),ox(*),oy(*),oz(*),ax(*),ay(*),az(*),body_grad(*)
      real(c_double)::xc(*),yc(*),zc(*),dx(*),dy(*),dz(*),xlo(*),xhi(*),ylo(*),yhi(*),zlo(*),zhi(*)
      real(c_double)::bounds(*),gx(*),gy(*),gz(*),vortex(*),advection(*),totals(*)
      integer(c_int)::solid(*)
      integer(c_int),value::nx,ny,nz,west,east,south,north
    end function
  end interface

contains
  ! The reference's use_openacc=.false. runs its host loops
  ! (src/fpm_stairstep_mpi_3d.f90:124,440); this library has only the GPU
  ! backend, so the call runs on the GPU and the first such call says so.
  subroutine backend_notice(use_openacc,routine)
    logical,intent(in)::use_openacc
    character(*),intent(in)::routine
    if(use_openacc.or.backend_noticed)return
    backend_noticed=.true.
    write(error_unit,'(3a)')'fpm_stairstep_mpi_3d: ',routine, &
         ': use_openacc=.false. requested, but this library has only the GPU backend; '// &
         'the call runs on the GPU'
  end subroutine backend_notice

  ! src/fpm_stairstep_mpi_3d.f90:25-36
  subroutine exchange_solid_halo_3d(solid,nx,ny,nz,west,east,south,north,comm)
    integer,intent(in)::nx,ny,nz,west,east,south,north,comm
    integer,intent(inout)::solid(0:nx+1,0:ny+1,nz)
    if(comm==comm)continue
    call stop_on(c_exchange_solid_halo(solid,nx,ny,nz,west,east,south,north))
  end subroutine exchange_solid_halo_3d

  ! src/fpm_stairstep_mpi_3d.f90:41-92
  subroutine enforce_stairstep_compatibility_3d(rhs,solid,body_grad,xlo,xhi,ylo,yhi,zlo,zhi,dx,dy,dz, &
       nx,ny,nz,west,east,south,north,comm,correction,imbalance)
    integer,intent(in)::nx,ny,nz,west,east,south,north,comm,solid(0:nx+1,0:ny+1,nz)
    real(real64),intent(in)::rhs(nx,ny,nz),body_grad(3),dx(nx),dy(ny),dz(nz)
    real(real64),intent(inout)::xlo(ny,nz),xhi(ny,nz),ylo(nx,nz),yhi(nx,nz),zlo(nx,ny),zhi(nx,ny)
    real(real64),intent(out)::correction,imbalance
    if(comm==comm)continue
    call stop_on(c_compat(rhs,solid,body_grad,xlo,xhi,ylo,yhi,zlo,zhi,dx,dy,dz,nx,ny,nz, &
         west,east,south,north,correction,imbalance))
  end subroutine enforce_stairstep_compatibility_3d

  ! src/fpm_stairstep_mpi_3d.f90:106-330
  subroutine solve_stairstep_poisson_3d(phi,rhs,solid,body_grad,xlo,xhi,ylo,yhi,zlo,zhi,xc,yc,zc,dx,dy,dz, &
       nx,ny,nz,west,east,south,north,comm,tolerance,max_sweeps,check_interval,sweeps,residual,use_openacc,method)
    integer,intent(in)::nx,ny,nz,west,east,south,north,comm,max_sweeps,check_interval
    integer,intent(in)::solid(0:nx+1,0:ny+1,nz)
    real(real64),intent(inout)::phi(0:nx+1,0:ny+1,nz)
    real(real64),intent(in)::rhs(nx,ny,nz),body_grad(3)
    real(real64),intent(in)::xlo(ny,nz),xhi(ny,nz),ylo(nx,nz),yhi(nx,nz),zlo(nx,ny),zhi(nx,ny)
    real(real64),intent(in)::xc(0:nx+1),yc(0:ny+1),zc(nz),dx(nx),dy(ny),dz(nz),tolerance
    integer,intent(out)::sweeps
    real(real64),intent(out)::residual
    logical,intent(in),optional::use_openacc
    character(*),intent(in),optional::method
    integer(c_int)::m
    if(comm==comm)continue
    if(present(use_openacc))call backend_notice(use_openacc,'solve_stairstep_poisson_3d')
    m=0
    if(present(method))then
      ! 'pcg-mixed' is this library's FP32-inner / FP64-outer PCG, 'fcg-mg'
      ! its flexible CG with an FP32 multigrid preconditioner.
      if(method/='pcg'.and.method/='jacobi'.and.method/='pcg-mixed'.and.method/='fcg-mg') &
           error stop 'stair-step method must be pcg or jacobi'
      if(method=='jacobi')m=1
      if(method=='pcg-mixed')m=2
      if(method=='fcg-mg')m=4
    endif
    call stop_on(c_solve(phi,rhs,solid,body_grad,xlo,xhi,ylo,yhi,zlo,zhi,xc,yc,zc,dx,dy,dz, &
         nx,ny,nz,west,east,south,north,tolerance,max_sweeps,check_interval,m,sweeps,residual))
  end subroutine solve_stairstep_poisson_3d

  ! src/fpm_stairstep_mpi_3d.f90:403-419
  function stairstep_residual_3d(phi,rhs,solid,body_grad,xlo,xhi,ylo,yhi,zlo,zhi,xc,yc,zc,dx,dy,dz, &
       nx,ny,nz,west,east,south,north,comm) result(value)
    integer,intent(in)::nx,ny,nz,west,east,south,north,comm,solid(0:nx+1,0:ny+1,nz)
    real(real64),intent(in)::phi(0:nx+1,0:ny+1,nz),rhs(nx,ny,nz),body_grad(3)
    real(real64),intent(in)::xlo(ny,nz),xhi(ny,nz),ylo(nx,nz),yhi(nx,nz),zlo(nx,ny),zhi(nx,ny)
    real(real64),intent(in)::xc(0:nx+1),yc(0:ny+1),zc(nz),dx(nx),dy(ny),dz(nz)
    real(real64)::value
    if(comm==comm)continue
    call stop_on(c_residual(phi,rhs,solid,body_grad,xlo,xhi,ylo,yhi,zlo,zhi,xc,yc,zc,dx,dy,dz, &
         nx,ny,nz,west,east,south,north,value))
  end function stairstep_residual_3d

  ! src/fpm_stairstep_mpi_3d.f90:426-469
  subroutine compute_stairstep_fpm_terms_3d(phi,u,v,w,ox,oy,oz,ax,ay,az,solid,body_grad, &
       xc,yc,zc,dx,dy,dz,xlo,xhi,ylo,yhi,zlo,zhi,bounds,gx,gy,gz,vortex,advection,totals, &
       nx,ny,nz,west,east,south,north,comm,use_openacc)
    integer,intent(in)::nx,ny,nz,west,east,south,north,comm,solid(0:nx+1,0:ny+1,nz)
    real(real64),intent(in)::phi(0:nx+1,0:ny+1,nz),u(nx,ny,nz),v(nx,ny,nz),w(nx,ny,nz)
    real(real64),intent(in)::ox(nx,ny,nz),oy(nx,ny,nz),oz(nx,ny,nz),ax(nx,ny,nz),ay(nx,ny,nz),az(nx,ny,nz)
    real(real64),intent(in)::body_grad(3),xc(0:nx+1),yc(0:ny+1),zc(nz),dx(nx),dy(ny),dz(nz)
    real(real64),intent(in)::xlo(ny,nz),xhi(ny,nz),ylo(nx,nz),yhi(nx,nz),zlo(nx,ny),zhi(nx,ny),bounds(6)
    real(real64),intent(out)::gx(nx,ny,nz),gy(nx,ny,nz),gz(nx,ny,nz),vortex(nx,ny,nz),advection(nx,ny,nz)
    real(real64),intent(out)::totals(2)
    logical,intent(in),optional::use_openacc
    if(comm==comm)continue
    if(present(use_openacc))call backend_notice(use_openacc,'compute_stairstep_fpm_terms_3d')
    call stop_on(c_terms(phi,u,v,w,ox,oy,oz,ax,ay,az,solid,body_grad,xc,yc,zc,dx,dy,dz, &
         xlo,xhi,ylo,yhi,zlo,zhi,bounds,gx,gy,gz,vortex,advection,totals,nx,ny,nz, &
         west,east,south,north))
  end subroutine compute_stairstep_fpm_terms_3d
end module fpm_stairstep_mpi_3d
