! ref_wrap.f90 — C-callable wrapper around routines of the reference itself,
! compiled from the reference's own sources (see oracle/Makefile) into
! oracle/_ref/libvlref.so.  Test infrastructure only: it checks the C oracle.
!
! Wrapped reference routines:
!   heal_boxes             src/polygon_fracture.f90:805-1004
!   generate_trackers      src/polygon_fracture.f90:767-803
!   scanline_fracture      src/polygon_fracture.f90:361-610
!   calculate_union_area_fast  src/geometry.f90:1840-1927
!   BuildRTree + PerformMerge  src/rtree.f90:55-133, src/pnummerge.f90:279-385
!   BuildRTree + ComputeInteractionsGPU  src/interaction.f90:204-313 (built
!                          without an offload target: the loop runs on the host)
module vlref_wrap
  use iso_c_binding
  use iso_fortran_env, only: int64, real64
  use CommonModule
  use GeometryModule
  use PolygonFractureModule
  use RTReeBuilder
  use DataStructuresModule
  use PNumMergeModule
  use RTreeBuilderGPU
  implicit none
contains
  subroutine to_boxes(n, c, b)
    integer(c_int64_t), intent(in) :: n
    integer(c_int32_t), intent(in) :: c(4*n)
    type(Box), intent(out) :: b(n)
    integer(c_int64_t) :: i
    do i = 1, n
       b(i)%x1 = c(4*i-3); b(i)%y1 = c(4*i-2); b(i)%x2 = c(4*i-1); b(i)%y2 = c(4*i)
    end do
  end subroutine to_boxes

  subroutine from_boxes(n, b, cap, c, nout)
    integer(c_int64_t), intent(in) :: n, cap
    type(Box), intent(in) :: b(:)
    integer(c_int32_t), intent(out) :: c(4*cap)
    integer(c_int64_t), intent(out) :: nout
    integer(c_int64_t) :: i
    nout = n
    if (n > cap) return
    do i = 1, n
       c(4*i-3) = b(i)%x1; c(4*i-2) = b(i)%y1; c(4*i-1) = b(i)%x2; c(4*i) = b(i)%y2
    end do
  end subroutine from_boxes

  ! heal_boxes on n boxes; writes up to cap boxes, returns the true count in nout.
  subroutine ref_heal_boxes(n, cin, cap, cout, nout) bind(C, name="ref_heal_boxes")
    integer(c_int64_t), value :: n, cap
    integer(c_int32_t), intent(in) :: cin(4*n)
    integer(c_int32_t), intent(out) :: cout(4*cap)
    integer(c_int64_t), intent(out) :: nout
    type(Box), allocatable :: b(:)
    integer(kind=int64) :: nres
    allocate(b(n))
    call to_boxes(n, cin, b)
    call heal_boxes(int(n, int64), b, nres)
    call from_boxes(int(nres, c_int64_t), b, cap, cout, nout)
  end subroutine ref_heal_boxes

  ! bbox minus the union of the n boxes (generate_trackers + scanline_fracture).
  subroutine ref_complement(n, cin, bb, cap, cout, nout) bind(C, name="ref_complement")
    integer(c_int64_t), value :: n, cap
    integer(c_int32_t), intent(in) :: cin(4*n), bb(4)
    integer(c_int32_t), intent(out) :: cout(4*cap)
    integer(c_int64_t), intent(out) :: nout
    type(Box), allocatable :: b(:), res(:)
    type(Box) :: bbox
    type(XYTracker), allocatable :: trk(:)
    allocate(b(n))
    call to_boxes(n, cin, b)
    bbox%x1 = bb(1); bbox%y1 = bb(2); bbox%x2 = bb(3); bbox%y2 = bb(4)
    call generate_trackers(b, bbox, trk)
    call scanline_fracture(trk, res)
    call from_boxes(int(size(res), c_int64_t), res, cap, cout, nout)
  end subroutine ref_complement

  function ref_union_area(n, cin) result(area) bind(C, name="ref_union_area")
    integer(c_int64_t), value :: n
    integer(c_int32_t), intent(in) :: cin(4*n)
    integer(c_int64_t) :: area
    type(Box), allocatable :: b(:)
    allocate(b(n))
    call to_boxes(n, cin, b)
    area = calculate_union_area_fast(b)
  end function ref_union_area

  ! PerformMerge over the n boxes in their given order (n >= 1): labels(i) =
  ! the union-find's arr(i) after uf_fullreduce (0 for untouched boxes).
  subroutine ref_pnum(n, cin, labels, area, perim) bind(C, name="ref_pnum")
    integer(c_int64_t), value :: n
    integer(c_int32_t), intent(in) :: cin(4*n)
    integer(c_int64_t), intent(out) :: labels(n)
    real(c_double), intent(out) :: area, perim
    type(Box), allocatable :: b(:)
    type(RTreeNode), allocatable :: nodes(:)
    type(UnionFind) :: uf
    integer(kind=int64) :: root
    real(kind=real64) :: a, p
    allocate(b(n))
    call to_boxes(n, cin, b)
    allocate(nodes(CalculateTotalNodes(int(n, int64), K_LEAF_CAPACITY)))
    root = 0
    call BuildRTree(b, K_LEAF_CAPACITY, nodes, root)
    call PerformMerge(uf, b, K_LEAF_CAPACITY, nodes, root, a, p)
    labels = uf%arr(1:n)
    area = a
    perim = p
  end subroutine ref_pnum

  ! ComputeInteractionsGPU's count of positive-area pairs (n >= 1).
  function ref_interactions(n, cin) result(cnt) bind(C, name="ref_interactions")
    integer(c_int64_t), value :: n
    integer(c_int32_t), intent(in) :: cin(4*n)
    integer(c_int64_t) :: cnt
    type(Box), allocatable :: b(:)
    type(RTreeNode), allocatable :: nodes(:)
    integer(kind=int64) :: root, nn, k
    allocate(b(n))
    call to_boxes(n, cin, b)
    nn = CalculateTotalNodes(int(n, int64), K_LEAF_CAPACITY)
    allocate(nodes(nn))
    root = 0
    call BuildRTree(b, K_LEAF_CAPACITY, nodes, root)
    call ComputeInteractionsGPU(nodes, nn, b, int(n, int64), root, k)
    cnt = k
  end function ref_interactions
end module vlref_wrap
