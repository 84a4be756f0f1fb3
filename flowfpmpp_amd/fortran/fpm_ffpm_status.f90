module fpm_ffpm_status
  ! Turns a nonzero libffpm status into error stop with the library's message
  ! (the reference's own text for the checks the reference makes).
  use,intrinsic :: iso_c_binding
  implicit none
  private
  public::stop_on
  interface
    function ffpm_last_error() bind(C,name='ffpm_last_error') result(p)
      import::c_ptr
      type(c_ptr)::p
    end function
    integer(c_size_t) function c_strlen(s) bind(C,name='strlen')
      import::c_ptr,c_size_t
      type(c_ptr),value::s
    end function
  end interface
contains
  subroutine stop_on(status)
    integer(c_int),intent(in)::status
    character(kind=c_char),pointer::chars(:)
    character(len=:),allocatable::msg
    type(c_ptr)::p
    integer::n,i
    if(status==0)return
    p=ffpm_last_error()
    n=int(c_strlen(p))
    call c_f_pointer(p,chars,[n])
    allocate(character(len=n)::msg)
    do i=1,n
      msg(i:i)=chars(i)
    enddo
    error stop msg
  end subroutine stop_on
end module fpm_ffpm_status
