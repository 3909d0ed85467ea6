! Reading the reference's configuration files from Fortran: iso_c_binding
! interfaces to the reader in libchswe.so (include/ch_swe_config.h) and a
! config type with the reference's config%get(value, key, default) pattern
! (generic_config_t, src/stuff/config/abstract_generic_config_mod.f90;
! parse_config_file, src/stuff/config/config_tools_mod.f90 L12-26).
! Errors stop the program through swe_hip_abort, as the reference's
! parcomm%abort does.
module swe_hip_cfg_mod

    use iso_c_binding, only : c_int, c_double, c_ptr, c_char, c_null_char, c_null_ptr, c_associated
    use swe_hip_capi_mod, only : swe_error_message
    use swe_hip_mod, only : swe_hip_abort

    implicit none

    private
    public :: config_hip_t, parse_config_file

    type :: config_hip_t
        type(c_ptr) :: h = c_null_ptr
    contains
        procedure :: has => cfg_has
        procedure :: get_real => cfg_get_real
        procedure :: get_int => cfg_get_int
        procedure :: get_logical => cfg_get_logical
        procedure :: get_str => cfg_get_str
        procedure :: destroy => cfg_destroy
    end type config_hip_t

    interface
        integer(c_int) function swe_cfg_read(path, cfg) bind(C, name="swe_cfg_read")
            import :: c_int, c_ptr, c_char
            character(kind=c_char), intent(in) :: path(*)
            type(c_ptr), intent(out) :: cfg
        end function
        integer(c_int) function swe_cfg_destroy(cfg) bind(C, name="swe_cfg_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: cfg
        end function
        integer(c_int) function swe_cfg_has(cfg, key) bind(C, name="swe_cfg_has")
            import :: c_int, c_ptr, c_char
            type(c_ptr), value :: cfg
            character(kind=c_char), intent(in) :: key(*)
        end function
        integer(c_int) function swe_cfg_get_real(cfg, key, val) bind(C, name="swe_cfg_get_real")
            import :: c_int, c_ptr, c_char, c_double
            type(c_ptr), value :: cfg
            character(kind=c_char), intent(in) :: key(*)
            real(c_double), intent(out) :: val
        end function
        integer(c_int) function swe_cfg_get_int(cfg, key, val) bind(C, name="swe_cfg_get_int")
            import :: c_int, c_ptr, c_char
            type(c_ptr), value :: cfg
            character(kind=c_char), intent(in) :: key(*)
            integer(c_int), intent(out) :: val
        end function
        integer(c_int) function swe_cfg_get_logical(cfg, key, val) bind(C, name="swe_cfg_get_logical")
            import :: c_int, c_ptr, c_char
            type(c_ptr), value :: cfg
            character(kind=c_char), intent(in) :: key(*)
            integer(c_int), intent(out) :: val
        end function
        integer(c_int) function swe_cfg_get_str(cfg, key, buf, buflen, slen) bind(C, name="swe_cfg_get_str")
            import :: c_int, c_ptr, c_char
            type(c_ptr), value :: cfg
            character(kind=c_char), intent(in) :: key(*)
            character(kind=c_char), intent(out) :: buf(*)
            integer(c_int), value :: buflen
            integer(c_int), intent(out) :: slen
        end function
    end interface

contains

    function parse_config_file(file_name) result(cfg)
        character(len=*), intent(in) :: file_name
        type(config_hip_t) :: cfg
        if (swe_cfg_read(file_name//c_null_char, cfg%h) /= 0) call swe_hip_abort(swe_error_message())
    end function parse_config_file

    logical function cfg_has(this, key)
        class(config_hip_t), intent(in) :: this
        character(len=*),    intent(in) :: key
        cfg_has = swe_cfg_has(this%h, key//c_null_char) /= 0
    end function cfg_has

    function cfg_get_real(this, key, default) result(v)
        class(config_hip_t),    intent(in) :: this
        character(len=*),       intent(in) :: key
        real(kind=8), optional, intent(in) :: default
        real(kind=8) :: v
        if (present(default) .and. .not. this%has(key)) then
            v = default
        else if (swe_cfg_get_real(this%h, key//c_null_char, v) /= 0) then
            call swe_hip_abort(swe_error_message())
        end if
    end function cfg_get_real

    function cfg_get_int(this, key, default) result(v)
        class(config_hip_t), intent(in) :: this
        character(len=*),    intent(in) :: key
        integer, optional,   intent(in) :: default
        integer :: v
        integer(c_int) :: iv
        if (present(default) .and. .not. this%has(key)) then
            v = default
        else if (swe_cfg_get_int(this%h, key//c_null_char, iv) /= 0) then
            call swe_hip_abort(swe_error_message())
        else
            v = iv
        end if
    end function cfg_get_int

    function cfg_get_logical(this, key, default) result(v)
        class(config_hip_t), intent(in) :: this
        character(len=*),    intent(in) :: key
        logical, optional,   intent(in) :: default
        logical :: v
        integer(c_int) :: iv
        if (present(default) .and. .not. this%has(key)) then
            v = default
        else if (swe_cfg_get_logical(this%h, key//c_null_char, iv) /= 0) then
            call swe_hip_abort(swe_error_message())
        else
            v = iv /= 0
        end if
    end function cfg_get_logical

    function cfg_get_str(this, key, default) result(v)
        class(config_hip_t),        intent(in) :: this
        character(len=*),           intent(in) :: key
        character(len=*), optional, intent(in) :: default
        character(len=:), allocatable :: v
        character(kind=c_char) :: buf(1024)
        integer(c_int) :: n
        integer :: i
        if (present(default) .and. .not. this%has(key)) then
            v = default
            return
        end if
        if (swe_cfg_get_str(this%h, key//c_null_char, buf, 1024, n) /= 0) call swe_hip_abort(swe_error_message())
        n = min(n, 1023)
        allocate(character(len=n) :: v)
        do i = 1, n
            v(i:i) = buf(i)
        end do
    end function cfg_get_str

    subroutine cfg_destroy(this)
        class(config_hip_t), intent(inout) :: this
        integer(c_int) :: rc
        if (c_associated(this%h)) rc = swe_cfg_destroy(this%h)
        this%h = c_null_ptr
    end subroutine cfg_destroy

end module swe_hip_cfg_mod
