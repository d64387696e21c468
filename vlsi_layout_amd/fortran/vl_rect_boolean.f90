! vl_rect_boolean.f90 — the reference's rectangle-fracture call surface
! (src/rectboolean.f90:55-181) on the gfx950 engine.
!
!   type(Box)         int64 rectangle [x1,x2) x [y1,y2) (src/rectboolean.f90:65-67)
!   type(Vector_Box)  growable box list with count, elements(:), push, append,
!                     clear (src/rectboolean.f90:78-86)
!   fracture_rectangle(target_rect, overlaps, out_fragments)
!                     appends target_rect minus the union of overlaps to
!                     out_fragments (src/rectboolean.f90:128-181)
!
! The difference is computed on the GPU by the engine's NOT operator
! (vl_boolean_op with the Boost NOT code), so the fragments are the canonical
! maximal-strip decomposition of the same point set: the reference's
! left / right / bottom / top cutting order gives another decomposition of it
! (for (0,0,10,10) minus (2,2,8,8) both give the same four boxes).  The engine
! works on int32 coordinates: coordinates outside that range stop the program,
! as the reference's own error paths do.  An empty target (x1 >= x2 or
! y1 >= y2) or empty overlap boxes stand for empty point sets.
module vl_rect_boolean
  use iso_c_binding
  use iso_fortran_env, only: int32, int64
  use vl_layer_ops, only: vl_engine_handle, vl_check
  implicit none
  private

  public :: Box, Vector_Box, fracture_rectangle

  integer, parameter :: K_COORDINATE_KIND = int64

  type :: Box
     integer(kind=K_COORDINATE_KIND) :: x1 = 0, y1 = 0, x2 = 0, y2 = 0
  end type Box

  type :: Vector_Box
     integer :: count = 0
     integer :: capacity = 0
     type(Box), allocatable :: elements(:)
   contains
     procedure :: push => vb_push
     procedure :: append => vb_append
     procedure :: clear => vb_clear
  end type Vector_Box

  type, bind(C) :: cbox
     integer(c_int32_t) :: x1, y1, x2, y2
  end type cbox

  integer(c_int), parameter :: VL_OK = 0, VL_ENOSPC = 2
  integer(c_int64_t), parameter :: OP_NOT = 3

  interface
     function c_boolean_op(e, a, an, b, bn, out, cap, out_n, ctrl, val) bind(C, name="vl_boolean_op") result(rc)
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: e, a, b, out
       integer(c_int64_t), value :: an, bn, cap, ctrl, val
       integer(c_int64_t) :: out_n
       integer(c_int) :: rc
     end function c_boolean_op
  end interface

contains

  subroutine vb_push(this, r)
    class(Vector_Box), intent(inout) :: this
    type(Box), intent(in) :: r
    type(Box), allocatable :: grown(:)
    if (.not. allocated(this%elements)) then
       this%capacity = max(this%capacity, 16)
       allocate(this%elements(this%capacity))
    else if (this%count >= size(this%elements)) then
       this%capacity = 2 * size(this%elements)
       allocate(grown(this%capacity))
       grown(1:this%count) = this%elements(1:this%count)
       call move_alloc(from=grown, to=this%elements)
    end if
    this%count = this%count + 1
    this%elements(this%count) = r
  end subroutine vb_push

  subroutine vb_append(this, other)
    class(Vector_Box), intent(inout) :: this
    class(Vector_Box), intent(in) :: other
    integer :: k
    do k = 1, other%count
       call this%push(other%elements(k))
    end do
  end subroutine vb_append

  subroutine vb_clear(this)
    class(Vector_Box), intent(inout) :: this
    this%count = 0
  end subroutine vb_clear

  logical function nonempty(b)
    type(Box), intent(in) :: b
    nonempty = b%x1 < b%x2 .and. b%y1 < b%y2
  end function nonempty

  function to_c(b) result(c)
    type(Box), intent(in) :: b
    type(cbox) :: c
    integer(int64), parameter :: lo = -huge(0_int32) - 1_int64, hi = huge(0_int32)
    if (min(b%x1, b%y1, b%x2, b%y2) < lo .or. max(b%x1, b%y1, b%x2, b%y2) > hi) &
         error stop "fracture_rectangle: coordinate outside the int32 range of the engine"
    c = cbox(int(b%x1, c_int32_t), int(b%y1, c_int32_t), int(b%x2, c_int32_t), int(b%y2, c_int32_t))
  end function to_c

  ! out_fragments += target_rect \ (union of overlaps%elements(1:count))
  subroutine fracture_rectangle(target_rect, overlaps, out_fragments)
    type(Box), intent(in) :: target_rect
    type(Vector_Box), intent(in) :: overlaps
    type(Vector_Box), intent(inout) :: out_fragments
    type(cbox), allocatable, target :: a(:), b(:), res(:)
    integer(c_int64_t) :: nb, cap, n
    integer(c_int) :: rc
    integer :: k
    if (.not. nonempty(target_rect)) return
    allocate(a(1), b(max(overlaps%count, 1)))
    a(1) = to_c(target_rect)
    nb = 0
    do k = 1, overlaps%count
       if (.not. nonempty(overlaps%elements(k))) cycle
       nb = nb + 1
       b(nb) = to_c(overlaps%elements(k))
    end do
    cap = 4 * nb + 4
    do
       if (allocated(res)) deallocate(res)
       allocate(res(cap))
       n = 0
       rc = c_boolean_op(vl_engine_handle(), c_loc(a(1)), 1_c_int64_t, c_loc(b(1)), nb, c_loc(res(1)), cap, n, &
            OP_NOT, 0_c_int64_t)
       if (rc /= VL_ENOSPC) exit
       cap = n
    end do
    call vl_check(rc, "fracture_rectangle")
    do k = 1, int(n)
       call out_fragments%push(Box(int(res(k)%x1, int64), int(res(k)%y1, int64), &
            int(res(k)%x2, int64), int(res(k)%y2, int64)))
    end do
  end subroutine fracture_rectangle

end module vl_rect_boolean
