! vl_layer_ops.f90 — the reference's Fortran layer-operator surface, backed by
! the gfx950 engine (libvlayout.so) through its C ABI (include/vl_api.h).
!
! Procedure names, argument order and meaning follow the reference:
!   CalculateAND, CalculateFrameNOT, CalculateNOT, CalculateOR, CalculateFrameAND,
!   CalculateGROWLayer, CalculateBoostOperation, MergeHealLayer
!                                   src/design.f90:20-31, interfaces :154-313
!   CalculateSingleLayerAND, CreateGrid, CreateEXTENT
!                                   src/design_impl.f90:385-509, :159-293
!   merge_boxes_using_scanline      src/boxmerge.f90:273-304
!   PreprocessLayer, PerformMergeLayer, CertifyHealing, NeedsSorting/NeedsPNum/
!   NeedsRTree/NeedsHealing, LAYER_STATE_*, Layer, Design
!                                   src/design.f90:33-103, :316-337, :937-1083,
!                                   src/design_impl.f90:22-56
!   heal_boxes                      src/polygon_fracture.f90:805-808
!   sort_event_trackers             src/event_sorter_apif.f90:30-43
!   SortBoxesDirect/SortBoxesIndirect  src/gpu_morton_sort.f90:41-77
!   K_BOOST_CONTROL_* codes         src/commonmodule.f90:12-14
! The Box type is the reference's {x1,y1,x2,y2} int32 rectangle made
! interoperable (bind(C)), so a Layer's box array is passed to the C ABI as is.
!
! Layers live on the host (layer_boxes(1:n_used)) as in the reference, and
! additionally keep a device-resident copy: an operator uploads an input only
! when its device copy is stale, and the result stays on the device.  With
! vl_eager_host = .true. (the default) every result is also copied back into
! layer_boxes, which is the reference's contract; with .false. a caller
! running a chain of operators skips those copies and calls layer_to_host
! when it needs the boxes.  A caller that edits layer_boxes directly calls
! layer_boxes_changed afterwards.
!
! The reference stops with `error stop` on every failure; so does this module,
! after printing the engine's error text.  Results are the canonical
! maximal-strip decomposition (the point set the reference computes; box order
! may differ from the reference's).
module vl_layer_ops
  use iso_c_binding
  use iso_fortran_env, only: int8, int32, int64, real64, error_unit
  implicit none
  private

  public :: Box, XYTracker, Layer
  public :: CalculateAND, CalculateFrameNOT, CalculateNOT, CalculateOR, CalculateXOR
  public :: CalculateGROWLayer, CalculateBoostOperation, MergeHealLayer, heal_boxes
  public :: CalculateFrameAND, CalculateSingleLayerAND, CreateGrid, CreateEXTENT
  public :: merge_boxes_using_scanline, vl_check
  public :: layer_input_dev, layer_output_dev, layer_output_done
  public :: sort_event_trackers, SortBoxesDirect, SortBoxesIndirect
  public :: LayerAssertZero, LayerKleeArea, LayerXorEmpty, CopyLayer
  public :: layer_set_boxes, layer_to_host, layer_boxes_changed, layer_release
  public :: vl_fortran_init, vl_fortran_finalize, vl_fortran_sync, vl_engine_handle
  public :: K_BOOST_CONTROL_XOR, K_BOOST_CONTROL_OR, K_BOOST_CONTROL_AND, K_BOOST_CONTROL_NOT
  public :: K_BOOST_CONTROL_MERGE, K_BOOST_CONTROL_SIZE
  public :: LAYER_STATE_NONE, LAYER_STATE_HEAL, LAYER_STATE_SORT, LAYER_STATE_PNUM, LAYER_STATE_RTREE
  public :: LAYER_STATE_SLSORT, LAYER_STATE_FINAL, LAYER_STATE_MAX, LAYER_STATE_EVERYTHING
  public :: UnionFind, Design
  public :: NeedsSorting, NeedsPNum, NeedsRTree, NeedsHealing
  public :: PreprocessLayer, PerformMergeLayer, CertifyHealing, LayerInteractionCount

  integer(int64), parameter :: K_BOOST_CONTROL_XOR = 0, K_BOOST_CONTROL_OR = 1, &
       K_BOOST_CONTROL_AND = 2, K_BOOST_CONTROL_NOT = 3, K_BOOST_CONTROL_MERGE = 4, &
       K_BOOST_CONTROL_SIZE = 5
  ! Layer state bits (src/design.f90:37-48).  Here SORT means the boxes are in
  ! Morton order of their centres, and RTREE is set together with SORT: the
  ! engine's spatial index is the tile binning each operator builds on the
  ! device, so no tree is kept on the layer.
  integer(c_int), parameter :: LAYER_STATE_NONE = 0, LAYER_STATE_HEAL = 1, LAYER_STATE_SORT = 2, &
       LAYER_STATE_PNUM = 4, LAYER_STATE_RTREE = 8, LAYER_STATE_SLSORT = 16, LAYER_STATE_FINAL = 32, &
       LAYER_STATE_MAX = 64, LAYER_STATE_EVERYTHING = LAYER_STATE_MAX - 1

  type, bind(C) :: Box
     integer(c_int32_t) :: x1 = 0, y1 = 0, x2 = 0, y2 = 0
  end type Box

  ! 12-byte event record of src/commonmodule.f90:30-34 (int32 X, Y; int8 winding)
  type, bind(C) :: XYTracker
     integer(c_int32_t) :: X = 0, Y = 0
     integer(c_int8_t) :: polygonNumber = 0
  end type XYTracker

  ! Polygon numbers of a layer after PNUM (src/datastructures.f90:26-38):
  ! arr(i) = 0 when box i touches no other box, else its component's number.
  type :: UnionFind
     integer(int64), allocatable :: arr(:)
  end type UnionFind

  type :: Layer
     integer(int64) :: n_used = 0
     integer(int64) :: n_alloc = 0
     type(Box), allocatable :: layer_boxes(:)
     integer(c_int) :: layerState = LAYER_STATE_NONE
     type(UnionFind) :: pnumtable          ! valid with LAYER_STATE_PNUM
     real(real64) :: area = 0, perimeter = 0   ! PNUM's overlap area and touch perimeter
     character(len=:), allocatable :: fileName
     type(c_ptr) :: dev = c_null_ptr      ! the engine's layer handle
     logical :: dev_valid = .false.        ! the device copy holds the layer
     logical :: host_valid = .true.        ! layer_boxes(1:n_used) holds the layer
  end type Layer

  ! A named set of layers (src/design.f90:95-103); layers are looked up by
  ! name with design_layer_index.
  type :: Design
     type(Layer), allocatable :: layers(:)
     character(len=1024), allocatable :: layerNames(:)
     type(Box) :: DESIGN_EXTENT
     character(len=:), allocatable :: designName
     character(len=:), allocatable :: fileName
  end type Design
  public :: design_layer_index
  public :: RestoreSnapToLayer, SaveLayerToSnap, LoadKLBin, WriteKLBin, ClearLayer, PNumHealLayer
  public :: COMPRESSION_METHOD_SNAPPY, COMPRESSION_METHOD_ZLIB, COMPRESSION_METHOD_ZSTD

  ! Layer files (src/commonmodule.f90:15-21, src/box_codec.f90): the
  ! compression method of SNAP files written at flush (1 Snappy, the
  ! reference's default), and temporary_layers 2, under which a missing layer
  ! file reads as an empty layer (src/diskdesign_impl.f90:87-97).
  integer(c_int32_t), parameter :: COMPRESSION_METHOD_SNAPPY = 1, COMPRESSION_METHOD_ZLIB = 2, &
       COMPRESSION_METHOD_ZSTD = 3
  integer, public :: temporary_layers = 1
  integer(c_int32_t), public :: K_COMPRESSION_METHOD_TO_USE = COMPRESSION_METHOD_SNAPPY

  logical, public :: vl_eager_host = .true.
  integer(c_int), public :: vl_device = 0
  type(c_ptr), save :: engine = c_null_ptr

  ! C ABI (include/vl_api.h)
  integer(c_int), parameter :: VL_OK = 0
  interface
     function c_last_error() bind(C, name="vl_last_error") result(p)
       import :: c_ptr
       type(c_ptr) :: p
     end function c_last_error
     function c_engine_create(device, e) bind(C, name="vl_engine_create") result(rc)
       import :: c_int, c_ptr
       integer(c_int), value :: device
       type(c_ptr) :: e
       integer(c_int) :: rc
     end function c_engine_create
     function c_engine_destroy(e) bind(C, name="vl_engine_destroy") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: e
       integer(c_int) :: rc
     end function c_engine_destroy
     function c_engine_sync(e) bind(C, name="vl_engine_sync") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: e
       integer(c_int) :: rc
     end function c_engine_sync
     function c_layer_create(e, l) bind(C, name="vl_layer_create") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: e
       type(c_ptr) :: l
       integer(c_int) :: rc
     end function c_layer_create
     function c_layer_destroy(l) bind(C, name="vl_layer_destroy") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: l
       integer(c_int) :: rc
     end function c_layer_destroy
     function c_layer_upload(l, host, n) bind(C, name="vl_layer_upload") result(rc)
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: l, host
       integer(c_int64_t), value :: n
       integer(c_int) :: rc
     end function c_layer_upload
     function c_layer_download(l, host, cap, n) bind(C, name="vl_layer_download") result(rc)
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: l, host
       integer(c_int64_t), value :: cap
       integer(c_int64_t) :: n
       integer(c_int) :: rc
     end function c_layer_download
     function c_layer_size(l, n) bind(C, name="vl_layer_size") result(rc)
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: l
       integer(c_int64_t) :: n
       integer(c_int) :: rc
     end function c_layer_size
     function c_grow(e, a, east, north, west, south, c) bind(C, name="vl_grow") result(rc)
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: e, a, c
       integer(c_int32_t), value :: east, north, west, south
       integer(c_int) :: rc
     end function c_grow
     function c_size(e, a, d, c) bind(C, name="vl_size") result(rc)
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: e, a, c
       integer(c_int32_t), value :: d
       integer(c_int) :: rc
     end function c_size
     function c_merge(e, a, c) bind(C, name="vl_merge") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: e, a, c
       integer(c_int) :: rc
     end function c_merge
     function c_self_and(e, a, c) bind(C, name="vl_self_and") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: e, a, c
       integer(c_int) :: rc
     end function c_self_and
     function c_grid(e, a, rows, cols, overlap, c) bind(C, name="vl_grid") result(rc)
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: e, a, c
       integer(c_int32_t), value :: rows, cols, overlap
       integer(c_int) :: rc
     end function c_grid
     function c_extent(e, a, c) bind(C, name="vl_extent") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: e, a, c
       integer(c_int) :: rc
     end function c_extent
     function c_pnum(e, a, labels, ncomp, area, perim) bind(C, name="vl_pnum") result(rc)
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: e, a, labels
       integer(c_int64_t) :: ncomp, area, perim
       integer(c_int) :: rc
     end function c_pnum
     function c_interaction_count(e, a, cnt) bind(C, name="vl_interaction_count") result(rc)
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: e, a
       integer(c_int64_t) :: cnt
       integer(c_int) :: rc
     end function c_interaction_count
     function c_layer_device_ptr(l, dev) bind(C, name="vl_layer_device_ptr") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: l
       type(c_ptr) :: dev
       integer(c_int) :: rc
     end function c_layer_device_ptr
     function c_morton_sort_direct(e, boxes, n) bind(C, name="vl_morton_sort_direct") result(rc)
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: e, boxes
       integer(c_int64_t), value :: n
       integer(c_int) :: rc
     end function c_morton_sort_direct
     function c_canonicalize(e, a, c) bind(C, name="vl_canonicalize") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: e, a, c
       integer(c_int) :: rc
     end function c_canonicalize
     function c_klee_area(e, a, area) bind(C, name="vl_klee_area") result(rc)
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: e, a
       integer(c_int64_t) :: area
       integer(c_int) :: rc
     end function c_klee_area
     function c_layer_load_klbin(e, l, path) bind(C, name="vl_layer_load_klbin") result(rc)
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: e, l
       character(kind=c_char), intent(in) :: path(*)
       integer(c_int) :: rc
     end function c_layer_load_klbin
     function c_layer_save_klbin(e, l, path) bind(C, name="vl_layer_save_klbin") result(rc)
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: e, l
       character(kind=c_char), intent(in) :: path(*)
       integer(c_int) :: rc
     end function c_layer_save_klbin
     function c_layer_save_snap(e, l, path, method, props, pnum) bind(C, name="vl_layer_save_snap") result(rc)
       import :: c_int, c_ptr, c_char, c_int32_t
       type(c_ptr), value :: e, l, pnum
       character(kind=c_char), intent(in) :: path(*)
       integer(c_int32_t), value :: method, props
       integer(c_int) :: rc
     end function c_layer_save_snap
     function c_layer_load_snap(e, l, path, props, pnum, cap) bind(C, name="vl_layer_load_snap") result(rc)
       import :: c_int, c_ptr, c_char, c_int32_t, c_int64_t
       type(c_ptr), value :: e, l, pnum
       character(kind=c_char), intent(in) :: path(*)
       integer(c_int32_t) :: props
       integer(c_int64_t), value :: cap
       integer(c_int) :: rc
     end function c_layer_load_snap
     function c_snap_info(path, props, total, chunks, method) bind(C, name="vl_snap_info") result(rc)
       import :: c_int, c_char, c_int32_t, c_int64_t
       character(kind=c_char), intent(in) :: path(*)
       integer(c_int32_t) :: props, method
       integer(c_int64_t) :: total, chunks
       integer(c_int) :: rc
     end function c_snap_info
     function c_xor_empty(e, a, b, empty) bind(C, name="vl_xor_empty") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: e, a, b
       integer(c_int) :: empty
       integer(c_int) :: rc
     end function c_xor_empty
     subroutine device_event_sort(d_arr, n) bind(C, name="device_event_sort")
       import :: c_ptr, c_int64_t
       type(c_ptr), value :: d_arr
       integer(c_int64_t), value :: n
     end subroutine device_event_sort
     subroutine thrust_sort_boxes_direct(d_boxes, n) bind(C, name="thrust_sort_boxes_direct")
       import :: c_ptr, c_int64_t
       type(c_ptr), value :: d_boxes
       integer(c_int64_t), value :: n
     end subroutine thrust_sort_boxes_direct
     subroutine thrust_sort_boxes_indirect(d_boxes, d_indices, n) bind(C, name="thrust_sort_boxes_indirect")
       import :: c_ptr, c_int64_t
       type(c_ptr), value :: d_boxes, d_indices
       integer(c_int64_t), value :: n
     end subroutine thrust_sort_boxes_indirect
  end interface

  abstract interface
     function binop_fn(e, a, b, c) bind(C) result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: e, a, b, c
       integer(c_int) :: rc
     end function binop_fn
  end interface
  interface
     function c_and(e, a, b, c) bind(C, name="vl_and") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: e, a, b, c
       integer(c_int) :: rc
     end function c_and
     function c_not(e, a, b, c) bind(C, name="vl_not") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: e, a, b, c
       integer(c_int) :: rc
     end function c_not
     function c_or(e, a, b, c) bind(C, name="vl_or") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: e, a, b, c
       integer(c_int) :: rc
     end function c_or
     function c_xor(e, a, b, c) bind(C, name="vl_xor") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: e, a, b, c
       integer(c_int) :: rc
     end function c_xor
  end interface

contains

  ! ---- engine ---------------------------------------------------------------

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    type(c_ptr) :: p
    integer :: i, n
    if (rc == VL_OK) return
    p = c_last_error()
    write(error_unit, '(A,A,I0)', advance='no') trim(what), ': engine status ', rc
    if (c_associated(p)) then
       call c_f_pointer(p, msg, [4096])
       n = 0
       do i = 1, 4096
          if (msg(i) == c_null_char) exit
          n = i
       end do
       write(error_unit, '(A)', advance='no') ': '
       do i = 1, n
          write(error_unit, '(A)', advance='no') msg(i)
       end do
    end if
    write(error_unit, *)
    error stop "vl_layer_ops: engine call failed"
  end subroutine check

  ! The engine status check for modules built on this one.
  subroutine vl_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    call check(rc, what)
  end subroutine vl_check

  subroutine vl_fortran_init(device)
    integer, intent(in), optional :: device
    if (c_associated(engine)) return
    if (present(device)) vl_device = int(device, c_int)
    call check(c_engine_create(vl_device, engine), "vl_engine_create")
  end subroutine vl_fortran_init

  function vl_engine_handle() result(e)
    type(c_ptr) :: e
    call vl_fortran_init()
    e = engine
  end function vl_engine_handle

  subroutine vl_fortran_sync()
    if (c_associated(engine)) call check(c_engine_sync(engine), "vl_engine_sync")
  end subroutine vl_fortran_sync

  ! Destroys the engine.  Layers still holding device copies keep the engine
  ! alive until layer_release is called on them.
  subroutine vl_fortran_finalize()
    if (.not. c_associated(engine)) return
    call check(c_engine_destroy(engine), "vl_engine_destroy")
    engine = c_null_ptr
  end subroutine vl_fortran_finalize

  ! ---- layer residency ------------------------------------------------------

  subroutine ensure_handle(L)
    type(Layer), intent(inout) :: L
    call vl_fortran_init()
    if (.not. c_associated(L%dev)) call check(c_layer_create(engine, L%dev), "vl_layer_create")
  end subroutine ensure_handle

  subroutine to_device(L)
    type(Layer), intent(inout), target :: L
    call ensure_handle(L)
    if (L%dev_valid) return
    if (L%n_used > 0) then
       if (.not. allocated(L%layer_boxes)) error stop "vl_layer_ops: layer has n_used > 0 but no boxes"
       if (size(L%layer_boxes, kind=int64) < L%n_used) error stop "vl_layer_ops: n_used exceeds layer_boxes"
       call check(c_layer_upload(L%dev, c_loc(L%layer_boxes(1)), L%n_used), "vl_layer_upload")
    else
       call check(c_layer_upload(L%dev, c_null_ptr, 0_c_int64_t), "vl_layer_upload")
    end if
    L%dev_valid = .true.
  end subroutine to_device

  ! Copies the device copy into layer_boxes (grown when too small).
  subroutine layer_to_host(L)
    type(Layer), intent(inout), target :: L
    integer(c_int64_t) :: n
    if (L%host_valid) return
    if (.not. L%dev_valid) error stop "vl_layer_ops: layer has no valid copy"
    call check(c_layer_size(L%dev, n), "vl_layer_size")
    if (.not. allocated(L%layer_boxes)) then
       allocate(L%layer_boxes(max(n, 1_int64)))
    else if (size(L%layer_boxes, kind=int64) < n) then
       deallocate(L%layer_boxes)
       allocate(L%layer_boxes(n))
    end if
    L%n_alloc = size(L%layer_boxes, kind=int64)
    if (n > 0) call check(c_layer_download(L%dev, c_loc(L%layer_boxes(1)), L%n_alloc, n), "vl_layer_download")
    L%n_used = n
    L%host_valid = .true.
  end subroutine layer_to_host

  ! Replaces the layer's boxes with b(1:n) (host copy; uploaded on next use).
  subroutine layer_set_boxes(L, b, n)
    type(Layer), intent(inout) :: L
    type(Box), intent(in) :: b(:)
    integer(int64), intent(in) :: n
    if (allocated(L%layer_boxes)) deallocate(L%layer_boxes)
    allocate(L%layer_boxes(max(n, 1_int64)))
    if (n > 0) L%layer_boxes(1:n) = b(1:n)
    L%n_used = n
    L%n_alloc = size(L%layer_boxes, kind=int64)
    L%layerState = LAYER_STATE_NONE
    L%host_valid = .true.
    L%dev_valid = .false.
  end subroutine layer_set_boxes

  ! Call after editing layer_boxes or n_used directly.
  subroutine layer_boxes_changed(L)
    type(Layer), intent(inout) :: L
    L%host_valid = .true.
    L%dev_valid = .false.
    L%layerState = LAYER_STATE_NONE
  end subroutine layer_boxes_changed

  subroutine layer_release(L)
    type(Layer), intent(inout) :: L
    if (c_associated(L%dev)) call check(c_layer_destroy(L%dev), "vl_layer_destroy")
    L%dev = c_null_ptr
    L%dev_valid = .false.
    if (.not. L%host_valid) then
       L%n_used = 0
       L%host_valid = .true.
    end if
  end subroutine layer_release

  ! Device handle of an input the operator may not modify (intent(in)): the
  ! layer's own device copy when it is current, else a temporary upload of
  ! its boxes (released with release_input).  The layer itself is unchanged.
  subroutine input_handle(A, h, temp)
    type(Layer), intent(in), target :: A
    type(c_ptr), intent(out) :: h
    logical, intent(out) :: temp
    if (A%dev_valid .and. c_associated(A%dev)) then
       h = A%dev
       temp = .false.
       return
    end if
    call vl_fortran_init()
    h = c_null_ptr
    call check(c_layer_create(engine, h), "vl_layer_create")
    temp = .true.
    if (A%n_used > 0) then
       if (.not. allocated(A%layer_boxes)) error stop "vl_layer_ops: layer has n_used > 0 but no boxes"
       if (size(A%layer_boxes, kind=int64) < A%n_used) error stop "vl_layer_ops: n_used exceeds layer_boxes"
       call check(c_layer_upload(h, c_loc(A%layer_boxes(1)), A%n_used), "vl_layer_upload")
    else
       call check(c_layer_upload(h, c_null_ptr, 0_c_int64_t), "vl_layer_upload")
    end if
  end subroutine input_handle

  subroutine release_input(h, temp)
    type(c_ptr), intent(in) :: h
    logical, intent(in) :: temp
    if (temp) call check(c_layer_destroy(h), "vl_layer_destroy")
  end subroutine release_input

  ! The result of an operator is on the device; record it (and copy it back
  ! in eager mode).  Results are canonical, i.e. healed.
  subroutine finish_output(C)
    type(Layer), intent(inout) :: C
    integer(c_int64_t) :: n
    call check(c_layer_size(C%dev, n), "vl_layer_size")
    C%n_used = n
    C%dev_valid = .true.
    C%host_valid = .false.
    C%layerState = LAYER_STATE_HEAL
    if (vl_eager_host) call layer_to_host(C)
  end subroutine finish_output

  ! Deferred operators (a recorded program): the device handle of an input
  ! (uploaded when its device copy is stale) ...
  function layer_input_dev(L) result(h)
    type(Layer), intent(inout), target :: L
    type(c_ptr) :: h
    call to_device(L)
    h = L%dev
  end function layer_input_dev

  ! ... of a layer a deferred operator writes (its device copy is taken as
  ! current from here on: the program writes it before any of its readers) ...
  function layer_output_dev(L) result(h)
    type(Layer), intent(inout) :: L
    type(c_ptr) :: h
    call ensure_handle(L)
    L%dev_valid = .true.
    L%host_valid = .false.
    h = L%dev
  end function layer_output_dev

  ! ... and, once the program has run, the result recorded like any operator's.
  subroutine layer_output_done(L)
    type(Layer), intent(inout) :: L
    call finish_output(L)
  end subroutine layer_output_done

  subroutine binop(f, what, A, B, C)
    procedure(binop_fn) :: f
    character(len=*), intent(in) :: what
    type(Layer), intent(inout) :: A, B, C
    call to_device(A)
    call to_device(B)
    call ensure_handle(C)
    call check(f(engine, A%dev, B%dev, C%dev), what)
    call finish_output(C)
  end subroutine binop

  ! ---- operators (src/design.f90 surface) -----------------------------------

  ! C = A AND B  ('*', src/design.f90:1131-1235)
  subroutine CalculateAND(A, B, C)
    type(Layer), intent(inout) :: A, B, C
    call binop(c_and, "CalculateAND", A, B, C)
  end subroutine CalculateAND

  ! C = A \ B  ('~', src/design_impl.f90:1152-1164)
  subroutine CalculateFrameNOT(A, B, C)
    type(Layer), intent(inout) :: A, B, C
    call binop(c_not, "CalculateFrameNOT", A, B, C)
  end subroutine CalculateFrameNOT

  ! C = A \ B  ('@', src/design.f90:1238-1268: complement of B in MBR(A), AND A)
  subroutine CalculateNOT(A, B, C)
    type(Layer), intent(inout) :: A, B, C
    call binop(c_not, "CalculateNOT", A, B, C)
  end subroutine CalculateNOT

  ! C = A OR B  (OR2, src/design.f90:1085-1103)
  subroutine CalculateOR(A, B, C)
    type(Layer), intent(inout) :: A, B, C
    call binop(c_or, "CalculateOR", A, B, C)
  end subroutine CalculateOR

  ! C = A XOR B  (XOR, src/design_impl.f90:961-991)
  subroutine CalculateXOR(A, B, C)
    type(Layer), intent(inout) :: A, B, C
    call binop(c_xor, "CalculateXOR", A, B, C)
  end subroutine CalculateXOR

  ! C = A grown by ivar = (E, N, W, S)  (src/design_impl.f90:571-586; A is
  ! intent(in) as in src/design.f90:243)
  subroutine CalculateGROWLayer(A, C, ivar)
    type(Layer), intent(in), target :: A
    type(Layer), intent(inout) :: C
    integer(int32), intent(in) :: ivar(4)
    type(c_ptr) :: h
    logical :: temp
    call input_handle(A, h, temp)
    call ensure_handle(C)
    call check(c_grow(engine, h, ivar(1), ivar(2), ivar(3), ivar(4), C%dev), "CalculateGROWLayer")
    call release_input(h, temp)
    call finish_output(C)
  end subroutine CalculateGROWLayer

  ! C = A AND B  (AND2, src/design.f90:296-299): the same point set as '*'.
  subroutine CalculateFrameAND(A, B, C)
    type(Layer), intent(inout) :: A, B, C
    call binop(c_and, "CalculateFrameAND", A, B, C)
  end subroutine CalculateFrameAND

  ! C = the points covered by at least two boxes of A (the one-input AND,
  ! src/design_impl.f90:385-509, src/controlparser.f90:545-549).
  subroutine CalculateSingleLayerAND(input_layer_A, output_layer)
    type(Layer), intent(inout) :: input_layer_A, output_layer
    call to_device(input_layer_A)
    call ensure_handle(output_layer)
    call check(c_self_and(engine, input_layer_A%dev, output_layer%dev), "CalculateSingleLayerAND")
    call finish_output(output_layer)
  end subroutine CalculateSingleLayerAND

  ! output_layer = rows x cols sub-boxes of MBR(input_layer), each padded by
  ! GridOverlap and clamped to the MBR, row-major (src/design_impl.f90:159-275).
  ! The boxes are kept as they are (the reference marks them healed so that
  ! they are not merged).
  subroutine CreateGrid(input_layer, output_layer, rows, cols, GridOverlap)
    type(Layer), intent(in), target :: input_layer
    type(Layer), intent(inout) :: output_layer
    integer, intent(in) :: rows, cols
    integer, intent(in) :: GridOverlap
    type(c_ptr) :: h
    logical :: temp
    if (rows < 1 .or. cols < 1) error stop "CreateGrid: rows and cols must be positive"
    call input_handle(input_layer, h, temp)
    call ensure_handle(output_layer)
    call check(c_grid(engine, h, int(rows, c_int32_t), int(cols, c_int32_t), int(GridOverlap, c_int32_t), &
         output_layer%dev), "CreateGrid")
    call release_input(h, temp)
    call finish_output(output_layer)
  end subroutine CreateGrid

  ! output_layer = MBR(input_layer) as one box (src/design_impl.f90:277-293).
  subroutine CreateEXTENT(input_layer, output_layer)
    type(Layer), intent(in), target :: input_layer
    type(Layer), intent(inout) :: output_layer
    type(c_ptr) :: h
    logical :: temp
    call input_handle(input_layer, h, temp)
    call ensure_handle(output_layer)
    call check(c_extent(engine, h, output_layer%dev), "CreateEXTENT")
    call release_input(h, temp)
    call finish_output(output_layer)
  end subroutine CreateEXTENT

  ! Boost-coded operators (src/design_impl.f90:931-951, src/boost_polygon_api.cpp:97-156):
  ! XOR=0, OR=1, AND=2, NOT=3, MERGE=4 (A only), SIZE=5 (A only, by control_value).
  ! An unknown code gives an empty C, as the reference's C++ side does.
  subroutine CalculateBoostOperation(A, B, C, control_parameter, control_value)
    type(Layer), intent(inout) :: A, B, C
    integer(int64), intent(in) :: control_parameter, control_value
    select case (control_parameter)
    case (K_BOOST_CONTROL_XOR)
       call binop(c_xor, "CalculateBoostOperation(XOR)", A, B, C)
    case (K_BOOST_CONTROL_OR)
       call binop(c_or, "CalculateBoostOperation(OR)", A, B, C)
    case (K_BOOST_CONTROL_AND)
       call binop(c_and, "CalculateBoostOperation(AND)", A, B, C)
    case (K_BOOST_CONTROL_NOT)
       call binop(c_not, "CalculateBoostOperation(NOT)", A, B, C)
    case (K_BOOST_CONTROL_MERGE)
       call to_device(A)
       call ensure_handle(C)
       call check(c_merge(engine, A%dev, C%dev), "CalculateBoostOperation(MERGE)")
       call finish_output(C)
    case (K_BOOST_CONTROL_SIZE)
       if (control_value < -huge(0_int32) - 1_int64 .or. control_value > huge(0_int32)) &
            error stop "CalculateBoostOperation: SIZE amount out of int32 range"
       call to_device(A)
       call ensure_handle(C)
       call check(c_size(engine, A%dev, int(control_value, c_int32_t), C%dev), "CalculateBoostOperation(SIZE)")
       call finish_output(C)
    case default
       call ensure_handle(C)
       call check(c_layer_upload(C%dev, c_null_ptr, 0_c_int64_t), "vl_layer_upload")
       call finish_output(C)
    end select
  end subroutine CalculateBoostOperation

  ! L = union of L, healed in place (src/design_impl.f90:1318-1341)
  subroutine MergeHealLayer(L)
    type(Layer), intent(inout) :: L
    call to_device(L)
    call check(c_merge(engine, L%dev, L%dev), "MergeHealLayer")
    call finish_output(L)
  end subroutine MergeHealLayer

  ! ---- layer state (src/design.f90:316-337) ---------------------------------

  pure logical function NeedsSorting(L)
    type(Layer), intent(in) :: L
    NeedsSorting = iand(L%layerState, LAYER_STATE_SORT) == 0
  end function NeedsSorting

  pure logical function NeedsPNum(L)
    type(Layer), intent(in) :: L
    NeedsPNum = iand(L%layerState, LAYER_STATE_PNUM) == 0
  end function NeedsPNum

  pure logical function NeedsRTree(L)
    type(Layer), intent(in) :: L
    NeedsRTree = iand(L%layerState, LAYER_STATE_RTREE) == 0
  end function NeedsRTree

  pure logical function NeedsHealing(L)
    type(Layer), intent(in) :: L
    NeedsHealing = iand(L%layerState, LAYER_STATE_HEAL) == 0
  end function NeedsHealing

  ! Morton order of the box centres, on the device copy (SortBoxesDirect's
  ! order, src/gpu_morton_sort.f90:41-57).
  subroutine sort_layer(L)
    type(Layer), intent(inout) :: L
    type(c_ptr) :: d
    call to_device(L)
    if (L%n_used < 2) return
    call check(c_layer_device_ptr(L%dev, d), "vl_layer_device_ptr")
    call check(c_morton_sort_direct(engine, d, L%n_used), "PreprocessLayer (sort)")
    L%host_valid = .false.
    if (vl_eager_host) call layer_to_host(L)
  end subroutine sort_layer

  ! Polygon numbers of the boxes in their current order (PerformMerge,
  ! src/pnummerge.f90:279-385) into pnumtable, area and perimeter.
  subroutine pnum_layer(L)
    type(Layer), intent(inout) :: L
    integer(int64), allocatable, target :: labels(:)
    integer(c_int64_t) :: nc, ar, pe
    call to_device(L)
    allocate(labels(max(L%n_used, 1_int64)))
    labels = 0
    call check(c_pnum(engine, L%dev, c_loc(labels(1)), nc, ar, pe), "PerformMerge")
    if (allocated(L%pnumtable%arr)) deallocate(L%pnumtable%arr)
    allocate(L%pnumtable%arr(L%n_used))
    L%pnumtable%arr = labels(1:L%n_used)
    L%area = real(ar, real64)
    L%perimeter = real(pe, real64)
  end subroutine pnum_layer

  ! PreprocessLayer (src/design.f90:937-1083, the PreprocessLayerSL form):
  ! heals a layer whose HEAL bit is clear, sorts it (SORT and RTREE), and
  ! numbers its polygons (PNUM).  An empty layer gets every state bit below
  ! SLSORT's (31), as in the reference.
  subroutine PreprocessLayer(L)
    type(Layer), intent(inout) :: L
    if (L%n_used == 0) then
       L%layerState = 31
       return
    end if
    if (NeedsHealing(L)) call MergeHealLayer(L)
    if (NeedsSorting(L) .or. NeedsRTree(L)) then
       call sort_layer(L)
       L%layerState = ior(L%layerState, ior(LAYER_STATE_SORT, LAYER_STATE_RTREE))
    end if
    if (NeedsPNum(L)) then
       call pnum_layer(L)
       L%layerState = ior(L%layerState, LAYER_STATE_PNUM)
    end if
  end subroutine PreprocessLayer

  ! PerformMergeLayer (src/design_impl.f90:22-33): heal when needed, then
  ! number the polygons.
  subroutine PerformMergeLayer(L)
    type(Layer), intent(inout) :: L
    if (NeedsHealing(L)) call MergeHealLayer(L)
    call pnum_layer(L)
    L%layerState = ior(L%layerState, LAYER_STATE_PNUM)
  end subroutine PerformMergeLayer

  ! Pairs of boxes overlapping with positive area (ComputeInteractionsGPU,
  ! src/interaction.f90:204-310).
  function LayerInteractionCount(L) result(n)
    type(Layer), intent(in) :: L
    integer(int64) :: n
    type(c_ptr) :: h
    logical :: temp
    integer(c_int64_t) :: cnt
    n = 0
    if (L%n_used < 2) return
    call input_handle(L, h, temp)
    call check(c_interaction_count(engine, h, cnt), "ComputeInteractionsGPU")
    call release_input(h, temp)
    n = cnt
  end function LayerInteractionCount

  ! CertifyHealing (src/design_impl.f90:35-56): true when fewer than two
  ! boxes, when the HEAL bit is set, or when no two boxes overlap with
  ! positive area.  The reference answers .false. for a layer without an
  ! R-tree; the engine needs none, so it counts the pairs instead.
  logical function CertifyHealing(L)
    type(Layer), intent(in) :: L
    CertifyHealing = .true.
    if (L%n_used < 2) return
    if (iand(L%layerState, LAYER_STATE_HEAL) /= 0) return
    CertifyHealing = LayerInteractionCount(L) == 0
  end function CertifyHealing

  ! 1-based index of the layer called name in d, 0 when there is none.
  integer function design_layer_index(d, name)
    type(Design), intent(in) :: d
    character(len=*), intent(in) :: name
    integer :: i
    design_layer_index = 0
    if (.not. allocated(d%layerNames)) return
    do i = 1, size(d%layerNames)
       if (trim(d%layerNames(i)) == trim(name)) then
          design_layer_index = i
          return
       end if
    end do
  end function design_layer_index

  ! boxes(1:n_in) -> their union as maximal strips in boxes(1:n_out), sorted by
  ! (x1, y1, x2, y2); boxes is reallocated when the result needs more room
  ! (src/polygon_fracture.f90:805-1004).
  subroutine heal_boxes(n_in, boxes, n_out)
    integer(int64), intent(in) :: n_in
    type(Box), allocatable, intent(inout), target :: boxes(:)
    integer(int64), intent(out) :: n_out
    type(c_ptr) :: a, c
    integer(c_int64_t) :: n
    n_out = 0
    if (n_in <= 0) return
    call vl_fortran_init()
    a = c_null_ptr
    c = c_null_ptr
    call check(c_layer_create(engine, a), "vl_layer_create")
    call check(c_layer_create(engine, c), "vl_layer_create")
    call check(c_layer_upload(a, c_loc(boxes(1)), n_in), "vl_layer_upload")
    call check(c_canonicalize(engine, a, c), "heal_boxes")
    call check(c_layer_size(c, n), "vl_layer_size")
    if (size(boxes, kind=int64) < n) then
       deallocate(boxes)
       allocate(boxes(n))
    end if
    if (n > 0) call check(c_layer_download(c, c_loc(boxes(1)), size(boxes, kind=c_int64_t), n), "vl_layer_download")
    n_out = n
    call check(c_layer_destroy(a), "vl_layer_destroy")
    call check(c_layer_destroy(c), "vl_layer_destroy")
  end subroutine heal_boxes

  ! boxes -> the union of boxes as maximal strips, sorted by (x1, y1, x2, y2),
  ! reallocated to the exact count (src/boxmerge.f90:273-304).  The reference
  ! joins touching boxes that form a rectangle and stops on overlapping ones;
  ! here any input is accepted and the result is the canonical decomposition
  ! of the same point set.
  subroutine merge_boxes_using_scanline(boxes)
    type(Box), allocatable, intent(inout), target :: boxes(:)
    type(Box), allocatable, target :: merged(:)
    type(c_ptr) :: a, c
    integer(c_int64_t) :: n
    if (.not. allocated(boxes)) return
    if (size(boxes) == 0) return
    call vl_fortran_init()
    a = c_null_ptr
    c = c_null_ptr
    call check(c_layer_create(engine, a), "vl_layer_create")
    call check(c_layer_create(engine, c), "vl_layer_create")
    call check(c_layer_upload(a, c_loc(boxes(1)), size(boxes, kind=c_int64_t)), "vl_layer_upload")
    call check(c_canonicalize(engine, a, c), "merge_boxes_using_scanline")
    call check(c_layer_size(c, n), "vl_layer_size")
    allocate(merged(n))
    if (n > 0) call check(c_layer_download(c, c_loc(merged(1)), n, n), "vl_layer_download")
    call move_alloc(from=merged, to=boxes)
    call check(c_layer_destroy(a), "vl_layer_destroy")
    call check(c_layer_destroy(c), "vl_layer_destroy")
  end subroutine merge_boxes_using_scanline

  ! ---- checks -----------------------------------------------------------------

  ! Number of boxes of the layer (ASSERT_ZERO, src/controlparser.f90:524-530).
  function LayerAssertZero(L) result(count)
    type(Layer), intent(in) :: L
    integer(int64) :: count
    count = L%n_used
  end function LayerAssertZero

  ! Exact union area (src/geometry.f90:1840-1927).
  function LayerKleeArea(L) result(area)
    type(Layer), intent(inout) :: L
    integer(int64) :: area
    integer(c_int64_t) :: a
    call to_device(L)
    call check(c_klee_area(engine, L%dev, a), "LayerKleeArea")
    area = a
  end function LayerKleeArea

  ! True when A XOR B is empty (src/boost_xor_checker.cpp:68-152).
  function LayerXorEmpty(A, B) result(empty)
    type(Layer), intent(inout) :: A, B
    logical :: empty
    integer(c_int) :: e
    call to_device(A)
    call to_device(B)
    call check(c_xor_empty(engine, A%dev, B%dev, e), "LayerXorEmpty")
    empty = e /= 0
  end function LayerXorEmpty

  ! lhs = rhs (COPY, src/controlparser.f90:555-559): the union of rhs.
  subroutine CopyLayer(lhs, rhs)
    type(Layer), intent(inout) :: lhs, rhs
    call to_device(rhs)
    call ensure_handle(lhs)
    call check(c_merge(engine, rhs%dev, lhs%dev), "CopyLayer")
    call finish_output(lhs)
  end subroutine CopyLayer

  ! ---- sorts (the GPU sort entry points of the reference) --------------------

  ! Stable sort of event trackers by (X, Y) (src/event_sorter_apif.f90:30-43).
  subroutine sort_event_trackers(arr)
    type(XYTracker), intent(inout), target :: arr(:)
    integer(int64) :: n
    n = size(arr, kind=int64)
    if (n <= 1) return
    call vl_fortran_init()
    call device_event_sort(c_loc(arr(1)), n)
  end subroutine sort_event_trackers

  ! Sort boxes by the Morton code of their centre (src/gpu_morton_sort.f90:41-55).
  subroutine SortBoxesDirect(Boxes, N)
    type(Box), intent(inout), target :: Boxes(:)
    integer(int64), intent(in) :: N
    if (N <= 1) return
    call vl_fortran_init()
    call thrust_sort_boxes_direct(c_loc(Boxes(1)), N)
  end subroutine SortBoxesDirect

  ! 1-based permutation that sorts Boxes by Morton code (src/gpu_morton_sort.f90:59-77).
  subroutine SortBoxesIndirect(Boxes, N, SortedIndices)
    type(Box), intent(in), target :: Boxes(:)
    integer(int64), intent(in) :: N
    integer(int64), intent(out), target :: SortedIndices(:)
    if (N <= 0) return
    if (N == 1) then
       SortedIndices(1) = 1
       return
    end if
    call vl_fortran_init()
    call thrust_sort_boxes_indirect(c_loc(Boxes(1)), c_loc(SortedIndices(1)), N)
  end subroutine SortBoxesIndirect

  ! The polygon pass of a design load (parseMagicLayoutFile,
  ! src/magparser.f90:358-418): number the polygons; a layer whose boxes
  ! overlap is healed and numbered again.  The layer ends HEAL and PNUM.
  subroutine PNumHealLayer(L)
    type(Layer), intent(inout) :: L
    if (L%n_used == 0) return
    call pnum_layer(L)
    if (L%area > 0) then
       if (CertifyHealing(L)) error stop "ERROR: INCORRECT GPU solution detected"
       call MergeHealLayer(L)
       call pnum_layer(L)
       if (L%area > 0) error stop 'UNION failed.'
    end if
    L%layerState = ior(L%layerState, ior(LAYER_STATE_PNUM, LAYER_STATE_HEAL))
  end subroutine PNumHealLayer

  ! ---- layer files ----------------------------------------------------------

  ! ClearLayer (src/design.f90:1279-1287): no boxes, every state bit set.
  subroutine ClearLayer(L)
    type(Layer), intent(inout) :: L
    call layer_release(L)
    if (allocated(L%layer_boxes)) deallocate(L%layer_boxes)
    if (allocated(L%pnumtable%arr)) deallocate(L%pnumtable%arr)
    L%n_used = 0
    L%n_alloc = 0
    L%host_valid = .true.
    L%dev_valid = .false.
    L%layerState = LAYER_STATE_EVERYTHING
  end subroutine ClearLayer

  ! RestoreSnapToLayer (src/diskdesign_impl.f90:83-114): a name holding
  ! ".bin" is a KLBIN file, anything else a SNAP file, read straight into the
  ! layer's device copy (SNAP chunks are decompressed on the host and their
  ! box codec decoded on the GPU).  With temporary_layers > 1 a missing file
  ! is an empty layer.  A SNAP file's state bits (masked to
  ! LAYER_STATE_EVERYTHING) become the layer's, and its polygon numbers, when
  ! present, fill pnumtable.  The reference then builds an R-tree (sorting
  ! the boxes); the engine needs none, so the file's order is kept.
  subroutine RestoreSnapToLayer(input_layer, snap_filename)
    type(Layer), intent(inout), target :: input_layer
    character(len=*), intent(in) :: snap_filename
    character(kind=c_char, len=:), allocatable :: cpath
    integer(c_int32_t) :: props, method
    integer(c_int64_t) :: total, nch, n
    logical :: exists
    inquire(file=trim(snap_filename), exist=exists)
    if (temporary_layers > 1 .and. .not. exists) then
       call ClearLayer(input_layer)
       return
    end if
    cpath = trim(snap_filename)//c_null_char
    call ClearLayer(input_layer)
    call ensure_handle(input_layer)
    if (index(snap_filename, '.bin') /= 0) then
       call check(c_layer_load_klbin(engine, input_layer%dev, cpath), "LoadKLBin")
       input_layer%layerState = LAYER_STATE_NONE
    else
       call check(c_snap_info(cpath, props, total, nch, method), "RestoreSnapToLayer")
       if (iand(props, LAYER_STATE_PNUM) /= 0 .and. total > 0) then
          allocate(input_layer%pnumtable%arr(total))
          call check(c_layer_load_snap(engine, input_layer%dev, cpath, props, c_loc(input_layer%pnumtable%arr(1)), &
               total), "RestoreSnapToLayer")
       else
          call check(c_layer_load_snap(engine, input_layer%dev, cpath, props, c_null_ptr, 0_c_int64_t), &
               "RestoreSnapToLayer")
       end if
       input_layer%layerState = iand(props, LAYER_STATE_EVERYTHING)
    end if
    call check(c_layer_size(input_layer%dev, n), "vl_layer_size")
    input_layer%n_used = n
    input_layer%dev_valid = .true.
    input_layer%host_valid = .false.
    if (vl_eager_host) call layer_to_host(input_layer)
  end subroutine RestoreSnapToLayer

  ! SaveLayerToSnap (src/diskdesign_impl.f90:59-81): the layer's boxes in a
  ! SNAP file; its state bits without LAYER_STATE_RTREE go to the header,
  ! its polygon numbers after it when LAYER_STATE_PNUM is set.  The
  ! reference sorts the layer for its R-tree first; here the stored order is
  ! written.
  subroutine SaveLayerToSnap(input_layer, snap_filename, method_to_use)
    type(Layer), intent(inout), target :: input_layer
    character(len=*), intent(in) :: snap_filename
    integer, intent(in) :: method_to_use
    type(c_ptr) :: pn
    integer(c_int32_t) :: props
    if (method_to_use /= COMPRESSION_METHOD_SNAPPY .and. method_to_use /= COMPRESSION_METHOD_ZLIB .and. &
         method_to_use /= COMPRESSION_METHOD_ZSTD) then
       write(error_unit, *) 'ERROR: unknown compression method: ', method_to_use, ' requested.'
       error stop
    end if
    call to_device(input_layer)
    props = iand(input_layer%layerState, not(LAYER_STATE_RTREE))
    pn = c_null_ptr
    if (iand(props, LAYER_STATE_PNUM) /= 0 .and. allocated(input_layer%pnumtable%arr) .and. input_layer%n_used > 0) then
       if (size(input_layer%pnumtable%arr, kind=int64) == input_layer%n_used) pn = c_loc(input_layer%pnumtable%arr(1))
    end if
    call check(c_layer_save_snap(engine, input_layer%dev, trim(snap_filename)//c_null_char, &
         int(method_to_use, c_int32_t), props, pn), "SaveLayerToSnap")
  end subroutine SaveLayerToSnap

  ! LoadKLBin (src/kldata.f90:14-63): boxes of a raw 16-byte-per-box file;
  ! an empty or missing file and an invalid box stop the program.
  subroutine LoadKLBin(fileName, boxes)
    character(len=*), intent(in) :: fileName
    type(Box), allocatable, intent(out) :: boxes(:)
    integer(int64) :: file_bytes, total_boxes, i
    integer :: u, ios
    inquire(file=trim(fileName), size=file_bytes)
    if (file_bytes <= 0) then
       write(error_unit, *) 'Error: KLBIN File ', trim(fileName), ' is empty or does not exist.'
       error stop
    end if
    total_boxes = file_bytes / 16
    allocate(boxes(total_boxes))
    open(newunit=u, file=trim(fileName), access='stream', form='unformatted', status='old', action='read', &
         iostat=ios)
    if (ios /= 0) error stop "Error: Could not open the binary file"
    if (total_boxes > 0) read(u, iostat=ios) boxes
    close(u)
    if (ios /= 0) error stop "Error occurred while reading the data array."
    do i = 1, total_boxes
       if (.not. (boxes(i)%x1 < boxes(i)%x2 .and. boxes(i)%y1 < boxes(i)%y2)) error stop "INVALID BOX detected on input"
    end do
  end subroutine LoadKLBin

  ! WriteKLBin (src/kldata.f90:64-97): boxes(1:total_boxes), each valid.
  subroutine WriteKLBin(fileName, boxes, total_boxes)
    character(len=*), intent(in) :: fileName
    type(Box), intent(in) :: boxes(:)
    integer(int64), intent(in) :: total_boxes
    integer(int64) :: i
    integer :: u, ios
    do i = 1, total_boxes
       if (.not. (boxes(i)%x1 < boxes(i)%x2 .and. boxes(i)%y1 < boxes(i)%y2)) then
          write(error_unit, *) 'Box: ', i, ' WRONG.'
          error stop "INVALID BOX"
       end if
    end do
    open(newunit=u, file=trim(fileName), access='stream', form='unformatted', status='replace', action='write', &
         iostat=ios)
    if (ios /= 0) error stop "Error: Could not open the binary file for writing."
    if (total_boxes > 0) write(u) boxes(1:total_boxes)
    close(u)
  end subroutine WriteKLBin

end module vl_layer_ops
