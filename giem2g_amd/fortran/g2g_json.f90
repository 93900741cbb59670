! g2g_json.f90 -- a small JSON reader for the solver deck.
!
! Reads a whole JSON document into a flat node table and looks values up by
! dotted paths ("Anomaly.Nx", "FGMRES.Tolerance").  It covers what the
! deck uses: objects, arrays, numbers, strings, true/false/null, and
! tolerates a trailing comma before ] or } (the reference's own decks have
! one, test/giem2g_input.json:44).  The reference reads the deck with the
! FSON library (giem2g.F90 via fson_get); this module is an independent
! reader with the same path semantics for the keys the driver needs.
module g2g_json
  implicit none
  private
  public :: json_doc, json_load, json_has, json_get_real, json_get_int, json_get_logical, &
            json_get_string, json_get_real_array, json_array_len, json_get_elem_real

  integer, parameter :: K_OBJ = 1, K_ARR = 2, K_NUM = 3, K_STR = 4, K_BOOL = 5, K_NULL = 6
  integer, parameter :: MAXKEY = 128

  type :: json_doc
     integer :: n = 0
     integer, allocatable :: kind(:), parent(:), first(:), next(:), slen(:), sstart(:)
     character(len=MAXKEY), allocatable :: key(:)
     double precision, allocatable :: num(:)
     logical, allocatable :: bool(:)
     character(len=:), allocatable :: text
  end type json_doc

contains

  subroutine grow(d)
    type(json_doc), intent(inout) :: d
    integer :: m
    integer, allocatable :: ti(:)
    character(len=MAXKEY), allocatable :: tk(:)
    double precision, allocatable :: tn(:)
    logical, allocatable :: tb(:)
    if (.not. allocated(d%kind)) then
       m = 256
       allocate(d%kind(m), d%parent(m), d%first(m), d%next(m), d%slen(m), d%sstart(m), &
                d%key(m), d%num(m), d%bool(m))
       return
    end if
    if (d%n < size(d%kind)) return
    m = 2 * size(d%kind)
    allocate(ti(m)); ti(1:d%n) = d%kind(1:d%n); call move_alloc(ti, d%kind)
    allocate(ti(m)); ti(1:d%n) = d%parent(1:d%n); call move_alloc(ti, d%parent)
    allocate(ti(m)); ti(1:d%n) = d%first(1:d%n); call move_alloc(ti, d%first)
    allocate(ti(m)); ti(1:d%n) = d%next(1:d%n); call move_alloc(ti, d%next)
    allocate(ti(m)); ti(1:d%n) = d%slen(1:d%n); call move_alloc(ti, d%slen)
    allocate(ti(m)); ti(1:d%n) = d%sstart(1:d%n); call move_alloc(ti, d%sstart)
    allocate(tk(m)); tk(1:d%n) = d%key(1:d%n); call move_alloc(tk, d%key)
    allocate(tn(m)); tn(1:d%n) = d%num(1:d%n); call move_alloc(tn, d%num)
    allocate(tb(m)); tb(1:d%n) = d%bool(1:d%n); call move_alloc(tb, d%bool)
  end subroutine grow

  integer function new_node(d, k, par, key) result(id)
    type(json_doc), intent(inout) :: d
    integer, intent(in) :: k, par
    character(len=*), intent(in) :: key
    integer :: c
    call grow(d)
    d%n = d%n + 1
    id = d%n
    d%kind(id) = k
    d%parent(id) = par
    d%first(id) = 0
    d%next(id) = 0
    d%slen(id) = 0
    d%sstart(id) = 0
    d%key(id) = key
    d%num(id) = 0d0
    d%bool(id) = .false.
    if (par > 0) then
       if (d%first(par) == 0) then
          d%first(par) = id
       else
          c = d%first(par)
          do while (d%next(c) /= 0)
             c = d%next(c)
          end do
          d%next(c) = id
       end if
    end if
  end function new_node

  subroutine skip_ws(d, p)
    type(json_doc), intent(in) :: d
    integer, intent(inout) :: p
    do while (p <= len(d%text))
       select case (iachar(d%text(p:p)))
       case (9, 10, 13, 32)
          p = p + 1
       case default
          return
       end select
    end do
  end subroutine skip_ws

  subroutine parse_string(d, p, s0, s1, ok)
    ! p at the opening quote; returns the content span [s0, s1]
    type(json_doc), intent(in) :: d
    integer, intent(inout) :: p
    integer, intent(out) :: s0, s1
    logical, intent(out) :: ok
    ok = .false.
    if (d%text(p:p) /= '"') return
    p = p + 1
    s0 = p
    do while (p <= len(d%text))
       if (d%text(p:p) == '\') then
          p = p + 2
          cycle
       end if
       if (d%text(p:p) == '"') then
          s1 = p - 1
          p = p + 1
          ok = .true.
          return
       end if
       p = p + 1
    end do
  end subroutine parse_string

  recursive subroutine parse_value(d, p, par, key, ok)
    type(json_doc), intent(inout) :: d
    integer, intent(inout) :: p
    integer, intent(in) :: par
    character(len=*), intent(in) :: key
    logical, intent(out) :: ok
    integer :: id, s0, s1, q, ios
    character :: c
    ok = .false.
    call skip_ws(d, p)
    if (p > len(d%text)) return
    c = d%text(p:p)
    if (c == '{') then
       id = new_node(d, K_OBJ, par, key)
       p = p + 1
       do
          call skip_ws(d, p)
          if (p > len(d%text)) return
          if (d%text(p:p) == '}') then
             p = p + 1
             exit
          end if
          call parse_string(d, p, s0, s1, ok)
          if (.not. ok) return
          call skip_ws(d, p)
          if (d%text(p:p) /= ':') then
             ok = .false.
             return
          end if
          p = p + 1
          call parse_value(d, p, id, d%text(s0:s1), ok)
          if (.not. ok) return
          call skip_ws(d, p)
          if (d%text(p:p) == ',') then
             p = p + 1
          else if (d%text(p:p) /= '}') then
             ok = .false.
             return
          end if
       end do
       ok = .true.
    else if (c == '[') then
       id = new_node(d, K_ARR, par, key)
       p = p + 1
       do
          call skip_ws(d, p)
          if (p > len(d%text)) return
          if (d%text(p:p) == ']') then
             p = p + 1
             exit
          end if
          call parse_value(d, p, id, '', ok)
          if (.not. ok) return
          call skip_ws(d, p)
          if (d%text(p:p) == ',') then
             p = p + 1
          else if (d%text(p:p) /= ']') then
             ok = .false.
             return
          end if
       end do
       ok = .true.
    else if (c == '"') then
       call parse_string(d, p, s0, s1, ok)
       if (.not. ok) return
       id = new_node(d, K_STR, par, key)
       d%sstart(id) = s0
       d%slen(id) = s1 - s0 + 1
    else if (d%text(p:min(p + 3, len(d%text))) == 'true') then
       id = new_node(d, K_BOOL, par, key)
       d%bool(id) = .true.
       p = p + 4
       ok = .true.
    else if (d%text(p:min(p + 4, len(d%text))) == 'false') then
       id = new_node(d, K_BOOL, par, key)
       p = p + 5
       ok = .true.
    else if (d%text(p:min(p + 3, len(d%text))) == 'null') then
       id = new_node(d, K_NULL, par, key)
       p = p + 4
       ok = .true.
    else
       q = p
       do while (q <= len(d%text))
          if (index('+-0123456789.eEdD', d%text(q:q)) == 0) exit
          q = q + 1
       end do
       if (q == p) return
       id = new_node(d, K_NUM, par, key)
       read(d%text(p:q - 1), *, iostat=ios) d%num(id)
       if (ios /= 0) return
       p = q
       ok = .true.
    end if
  end subroutine parse_value

  subroutine json_load(path, d, ok)
    character(len=*), intent(in) :: path
    type(json_doc), intent(out) :: d
    logical, intent(out) :: ok
    integer :: u, sz, ios, p
    ok = .false.
    open(newunit=u, file=path, access='stream', form='unformatted', status='old', iostat=ios)
    if (ios /= 0) return
    inquire(unit=u, size=sz)
    allocate(character(len=sz) :: d%text)
    read(u, iostat=ios) d%text
    close(u)
    if (ios /= 0) return
    p = 1
    call parse_value(d, p, 0, '', ok)
  end subroutine json_load

  integer function find(d, path) result(id)
    ! node of a dotted path below the root object, 0 when absent
    type(json_doc), intent(in) :: d
    character(len=*), intent(in) :: path
    integer :: a, b, c
    id = 0
    if (d%n == 0) return
    id = 1
    a = 1
    do while (a <= len_trim(path))
       b = index(path(a:), '.')
       if (b == 0) then
          b = len_trim(path) + 1
       else
          b = a + b - 1
       end if
       c = d%first(id)
       do while (c /= 0)
          if (trim(d%key(c)) == path(a:b - 1)) exit
          c = d%next(c)
       end do
       id = c
       if (id == 0) return
       a = b + 1
    end do
  end function find

  logical function json_has(d, path)
    type(json_doc), intent(in) :: d
    character(len=*), intent(in) :: path
    json_has = find(d, path) /= 0
  end function json_has

  double precision function json_get_real(d, path, default) result(v)
    type(json_doc), intent(in) :: d
    character(len=*), intent(in) :: path
    double precision, intent(in) :: default
    integer :: id
    v = default
    id = find(d, path)
    if (id /= 0) then
       if (d%kind(id) == K_NUM) v = d%num(id)
    end if
  end function json_get_real

  integer function json_get_int(d, path, default) result(v)
    type(json_doc), intent(in) :: d
    character(len=*), intent(in) :: path
    integer, intent(in) :: default
    v = nint(json_get_real(d, path, dble(default)))
  end function json_get_int

  logical function json_get_logical(d, path, default) result(v)
    type(json_doc), intent(in) :: d
    character(len=*), intent(in) :: path
    logical, intent(in) :: default
    integer :: id
    v = default
    id = find(d, path)
    if (id /= 0) then
       if (d%kind(id) == K_BOOL) v = d%bool(id)
    end if
  end function json_get_logical

  subroutine json_get_string(d, path, s, found)
    type(json_doc), intent(in) :: d
    character(len=*), intent(in) :: path
    character(len=*), intent(out) :: s
    logical, intent(out) :: found
    integer :: id
    s = ''
    found = .false.
    id = find(d, path)
    if (id == 0) return
    if (d%kind(id) /= K_STR) return
    if (d%slen(id) > 0) s = d%text(d%sstart(id):d%sstart(id) + d%slen(id) - 1)
    found = .true.
  end subroutine json_get_string

  integer function json_array_len(d, path) result(n)
    type(json_doc), intent(in) :: d
    character(len=*), intent(in) :: path
    integer :: id, c
    n = -1
    id = find(d, path)
    if (id == 0) return
    if (d%kind(id) /= K_ARR) return
    n = 0
    c = d%first(id)
    do while (c /= 0)
       n = n + 1
       c = d%next(c)
    end do
  end function json_array_len

  subroutine json_get_real_array(d, path, v, n)
    ! numbers of an array (n = count, -1 when absent); v must be large enough
    type(json_doc), intent(in) :: d
    character(len=*), intent(in) :: path
    double precision, intent(out) :: v(:)
    integer, intent(out) :: n
    integer :: id, c
    n = -1
    id = find(d, path)
    if (id == 0) return
    if (d%kind(id) /= K_ARR) return
    n = 0
    c = d%first(id)
    do while (c /= 0)
       n = n + 1
       if (n <= size(v)) v(n) = d%num(c)
       c = d%next(c)
    end do
  end subroutine json_get_real_array

  double precision function json_get_elem_real(d, path, i, field, default) result(v)
    ! number `field` of the i-th (1-based) object of the array at path
    ! ("Recievers", 2, "depth"); default when absent
    type(json_doc), intent(in) :: d
    character(len=*), intent(in) :: path, field
    integer, intent(in) :: i
    double precision, intent(in) :: default
    integer :: id, c, k
    v = default
    id = find(d, path)
    if (id == 0) return
    if (d%kind(id) /= K_ARR) return
    c = d%first(id)
    k = 1
    do while (c /= 0 .and. k < i)
       c = d%next(c)
       k = k + 1
    end do
    if (c == 0) return
    if (d%kind(c) /= K_OBJ) return
    c = d%first(c)
    do while (c /= 0)
       if (trim(d%key(c)) == field) then
          if (d%kind(c) == K_NUM) v = d%num(c)
          return
       end if
       c = d%next(c)
    end do
  end function json_get_elem_real

end module g2g_json
