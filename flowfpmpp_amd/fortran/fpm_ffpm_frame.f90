module fpm_ffpm_frame
  ! Keeps a caller's host arrays resident in HBM across calls of the drop-in
  ! modules, the role the caller's OpenACC data region plays for the
  ! reference (src/fpm_stairstep_mpi_3d.f90:145-148,
  ! tests/benchmark_stairstep_3d.f90:151):
  !
  !   type(ffpm_frame)::frame
  !   call ffpm_frame_begin(frame)                       ! !$acc data
  !   call ffpm_frame_bind(frame,rhs)                    ! copyin(rhs)
  !   call ffpm_frame_bind(frame,phi,copy_out=.true.)    ! copy(phi)
  !   ... compatibility, solve, terms: no array moves ...
  !   call ffpm_frame_end(frame)                         ! !$acc end data
  !
  ! While bound, the drop-in routines use each array's device copy in place;
  ! the host copy of an array is stale until ffpm_frame_end (copy_out arrays)
  ! or ffpm_frame_update_host. ffpm_frame_update_device uploads an array the
  ! host has changed. Library: libffpm (include/ffpm.h, ffpm_frame_*).
  use,intrinsic :: iso_fortran_env,only:real64,int64
  use,intrinsic :: iso_c_binding
  use fpm_ffpm_status,only:stop_on
  implicit none
  private
  public::ffpm_frame,ffpm_frame_begin,ffpm_frame_bind,ffpm_frame_update_device
  public::ffpm_frame_update_host,ffpm_frame_end,ffpm_transfer_counts

  type::ffpm_frame
    type(c_ptr)::handle=c_null_ptr
  end type ffpm_frame

  interface ffpm_frame_bind
    module procedure bind_r1,bind_r2,bind_r3,bind_i3
  end interface
  interface ffpm_frame_update_device
    module procedure upload_r1,upload_r2,upload_r3,upload_i3
  end interface
  interface ffpm_frame_update_host
    module procedure download_r1,download_r2,download_r3,download_i3
  end interface

  interface
    integer(c_int) function c_create(frame) bind(C,name='ffpm_frame_create')
      import::c_int,c_ptr
      type(c_ptr)::frame
    end function
    integer(c_int) function c_bind(frame,host,bytes,copy_out) bind(C,name='ffpm_frame_bind')
      import::c_int,c_ptr,c_size_t
      type(c_ptr),value::frame
      type(*),dimension(*)::host
      integer(c_size_t),value::bytes
      integer(c_int),value::copy_out
    end function
    integer(c_int) function c_upload(frame,host) bind(C,name='ffpm_frame_upload')
      import::c_int,c_ptr
      type(c_ptr),value::frame
      type(*),dimension(*)::host
    end function
    integer(c_int) function c_download(frame,host) bind(C,name='ffpm_frame_download')
      import::c_int,c_ptr
      type(c_ptr),value::frame
      type(*),dimension(*)::host
    end function
    integer(c_int) function c_destroy(frame) bind(C,name='ffpm_frame_destroy')
      import::c_int,c_ptr
      type(c_ptr),value::frame
    end function
    integer(c_int) function c_counts(counts) bind(C,name='ffpm_transfer_counts')
      import::c_int,c_long_long
      integer(c_long_long)::counts(4)
    end function
  end interface

contains
  subroutine ffpm_frame_begin(frame)
    type(ffpm_frame),intent(inout)::frame
    call stop_on(c_create(frame%handle))
  end subroutine ffpm_frame_begin

  subroutine ffpm_frame_end(frame)
    type(ffpm_frame),intent(inout)::frame
    call stop_on(c_destroy(frame%handle))
    frame%handle=c_null_ptr
  end subroutine ffpm_frame_end

  integer(c_int) function flag(copy_out)
    logical,intent(in),optional::copy_out
    flag=0
    if(present(copy_out))then;if(copy_out)flag=1;endif
  end function flag

  subroutine bind_r1(frame,a,copy_out)
    type(ffpm_frame),intent(in)::frame
    real(real64),intent(inout),contiguous::a(:)
    logical,intent(in),optional::copy_out
    call stop_on(c_bind(frame%handle,a,int(size(a,kind=int64)*8,c_size_t),flag(copy_out)))
  end subroutine bind_r1

  subroutine bind_r2(frame,a,copy_out)
    type(ffpm_frame),intent(in)::frame
    real(real64),intent(inout),contiguous::a(:,:)
    logical,intent(in),optional::copy_out
    call stop_on(c_bind(frame%handle,a,int(size(a,kind=int64)*8,c_size_t),flag(copy_out)))
  end subroutine bind_r2

  subroutine bind_r3(frame,a,copy_out)
    type(ffpm_frame),intent(in)::frame
    real(real64),intent(inout),contiguous::a(:,:,:)
    logical,intent(in),optional::copy_out
    call stop_on(c_bind(frame%handle,a,int(size(a,kind=int64)*8,c_size_t),flag(copy_out)))
  end subroutine bind_r3

  subroutine bind_i3(frame,a,copy_out)
    type(ffpm_frame),intent(in)::frame
    integer,intent(inout),contiguous::a(:,:,:)
    logical,intent(in),optional::copy_out
    call stop_on(c_bind(frame%handle,a,int(size(a,kind=int64)*(storage_size(a)/8),c_size_t), &
         flag(copy_out)))
  end subroutine bind_i3

  subroutine upload_r1(frame,a)
    type(ffpm_frame),intent(in)::frame
    real(real64),intent(in),contiguous::a(:)
    call stop_on(c_upload(frame%handle,a))
  end subroutine upload_r1

  subroutine upload_r2(frame,a)
    type(ffpm_frame),intent(in)::frame
    real(real64),intent(in),contiguous::a(:,:)
    call stop_on(c_upload(frame%handle,a))
  end subroutine upload_r2

  subroutine upload_r3(frame,a)
    type(ffpm_frame),intent(in)::frame
    real(real64),intent(in),contiguous::a(:,:,:)
    call stop_on(c_upload(frame%handle,a))
  end subroutine upload_r3

  subroutine upload_i3(frame,a)
    type(ffpm_frame),intent(in)::frame
    integer,intent(in),contiguous::a(:,:,:)
    call stop_on(c_upload(frame%handle,a))
  end subroutine upload_i3

  subroutine download_r1(frame,a)
    type(ffpm_frame),intent(in)::frame
    real(real64),intent(inout),contiguous::a(:)
    call stop_on(c_download(frame%handle,a))
  end subroutine download_r1

  subroutine download_r2(frame,a)
    type(ffpm_frame),intent(in)::frame
    real(real64),intent(inout),contiguous::a(:,:)
    call stop_on(c_download(frame%handle,a))
  end subroutine download_r2

  subroutine download_r3(frame,a)
    type(ffpm_frame),intent(in)::frame
    real(real64),intent(inout),contiguous::a(:,:,:)
    call stop_on(c_download(frame%handle,a))
  end subroutine download_r3

  subroutine download_i3(frame,a)
    type(ffpm_frame),intent(in)::frame
    integer,intent(inout),contiguous::a(:,:,:)
    call stop_on(c_download(frame%handle,a))
  end subroutine download_i3

  ! Host-device copies made by the library so far: uploads, bytes uploaded,
  ! downloads, bytes downloaded.
  subroutine ffpm_transfer_counts(counts)
    integer(int64),intent(out)::counts(4)
    integer(c_long_long)::c(4)
    call stop_on(c_counts(c))
    counts=int(c,int64)
  end subroutine ffpm_transfer_counts
end module fpm_ffpm_frame
