module fpm_trace
  ! Drop-in replacement of the reference module of the same name
  ! (src/fpm_trace.f90:1-14, NVTX build src/fpm_trace_nvtx.f90:1-16): named
  ! ranges on the profiler timeline, here roctx ranges that rocprofv3 shows
  ! with --marker-trace, through the C ABI of libffpm (include/ffpm.h).
  use,intrinsic :: iso_c_binding
  implicit none
  private
  public::trace_push,trace_pop

  interface
    integer(c_int) function c_trace_push(name) bind(C,name='ffpm_trace_push')
      import::c_int,c_char
      character(kind=c_char)::name(*)
    end function
    integer(c_int) function c_trace_pop() bind(C,name='ffpm_trace_pop')
      import::c_int
    end function
  end interface

contains
  subroutine trace_push(name)
    character(*),intent(in)::name
    integer(c_int)::status
    status=c_trace_push(name//c_null_char)
  end subroutine trace_push

  subroutine trace_pop()
    integer(c_int)::status
    status=c_trace_pop()
  end subroutine trace_pop
end module fpm_trace
