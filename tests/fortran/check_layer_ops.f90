! check_layer_ops.f90 — calls the Fortran call surface directly (test
! helper; run by tests/test_fortran.py on a GPU host).
!   check_layer_ops A.bin B.bin OUTDIR
! prints the known-answer results of fracture_rectangle and
! merge_boxes_using_scanline, and writes the results of CalculateFrameAND,
! CalculateSingleLayerAND, CreateGrid, CreateEXTENT and CalculateGROWLayer
! (called through an intent(in) argument) as KLBIN files in OUTDIR; then
! PreprocessLayer, PerformMergeLayer, CertifyHealing and a Design: the
! state bits and counts are printed, the boxes and polygon numbers written.
program check_layer_ops
  use iso_fortran_env, only: int64, int32
  use vl_layer_ops
  use vl_rect_boolean, only: RBox => Box, Vector_Box, fracture_rectangle
  implicit none
  character(len=512) :: fa, fb, dir
  type(Layer) :: A, B, C, P, Q, E
  type(Design) :: d
  type(Vector_Box) :: overlaps, frags
  type(Box), allocatable :: m(:)
  integer :: k

  call get_command_argument(1, fa)
  call get_command_argument(2, fb)
  call get_command_argument(3, dir)
  call vl_fortran_init()

  ! (0,0,10,10) minus (2,2,8,8): the four fragments around the hole
  call overlaps%push(RBox(2_int64, 2_int64, 8_int64, 8_int64))
  call fracture_rectangle(RBox(0_int64, 0_int64, 10_int64, 10_int64), overlaps, frags)
  print '(A,I0)', 'FRACTURE ', frags%count
  do k = 1, frags%count
     print '(A,4(1X,I0))', 'F', frags%elements(k)%x1, frags%elements(k)%y1, frags%elements(k)%x2, frags%elements(k)%y2
  end do
  ! an empty overlap list leaves the target; a covering one removes it
  call frags%clear()
  call overlaps%clear()
  call fracture_rectangle(RBox(-5_int64, -5_int64, 3_int64, 4_int64), overlaps, frags)
  call overlaps%push(RBox(-10_int64, -10_int64, 10_int64, 10_int64))
  call fracture_rectangle(RBox(0_int64, 0_int64, 1_int64, 1_int64), overlaps, frags)
  print '(A,I0,4(1X,I0))', 'FRACTURE2 ', frags%count, frags%elements(1)%x1, frags%elements(1)%y1, &
       frags%elements(1)%x2, frags%elements(1)%y2

  ! (0,0,2,2) + (2,0,4,2) -> (0,0,4,2)
  allocate(m(2))
  m(1) = Box(0, 0, 2, 2)
  m(2) = Box(2, 0, 4, 2)
  call merge_boxes_using_scanline(m)
  print '(A,I0,4(1X,I0))', 'MERGE ', size(m), m(1)%x1, m(1)%y1, m(1)%x2, m(1)%y2

  call load(A, fa)
  call load(B, fb)
  call CalculateFrameAND(A, B, C)
  call save(C, 'frame_and.bin')
  call CalculateSingleLayerAND(A, C)
  call save(C, 'self_and.bin')
  call CreateGrid(A, C, 3, 4, 2)
  call save(C, 'grid.bin')
  call CreateEXTENT(A, C)
  call save(C, 'extent.bin')
  call grow_of(A, C)
  call save(C, 'grow.bin')

  ! PreprocessLayer on the raw (overlapping) layer: heal, Morton sort, PNUM
  print '(A,L1,1X,I0)', 'CERTIFY_RAW ', CertifyHealing(A), LayerInteractionCount(A)
  call load(P, fa)
  call PreprocessLayer(P)
  print '(A,I0)', 'STATE ', P%layerState
  print '(A,L1,1X,L1,1X,L1,1X,L1)', 'NEEDS ', NeedsHealing(P), NeedsSorting(P), NeedsRTree(P), NeedsPNum(P)
  print '(A,L1,1X,I0)', 'CERTIFY_PRE ', CertifyHealing(P), LayerInteractionCount(P)
  print '(A,I0,1X,F0.1,1X,F0.1)', 'PNUM ', maxval(P%pnumtable%arr), P%area, P%perimeter
  call save(P, 'pre.bin')
  call save_labels(P, 'pre_labels.bin')
  ! PerformMergeLayer numbers the boxes of B after healing it
  call load(Q, fb)
  call PerformMergeLayer(Q)
  call save(Q, 'pml.bin')
  call save_labels(Q, 'pml_labels.bin')
  print '(A,I0)', 'PML_STATE ', Q%layerState
  ! an empty layer gets the state 31
  call layer_set_boxes(E, m, 0_int64)
  call PreprocessLayer(E)
  print '(A,I0)', 'EMPTY_STATE ', E%layerState
  ! a design holding two named layers
  allocate(d%layers(2), d%layerNames(2))
  d%layerNames(1) = 'metal1'
  d%layerNames(2) = 'via1'
  d%designName = 'check'
  print '(A,I0,1X,I0,1X,I0)', 'DESIGN ', design_layer_index(d, 'via1'), design_layer_index(d, 'metal1'), &
       design_layer_index(d, 'poly')
  print '(A)', 'CHECK DONE'

contains

  ! GROW of a layer the caller holds as intent(in) (src/design.f90:243)
  subroutine grow_of(src, dst)
    type(Layer), intent(in) :: src
    type(Layer), intent(inout) :: dst
    call CalculateGROWLayer(src, dst, [1_int32, 2_int32, 3_int32, 4_int32])
  end subroutine grow_of

  subroutine load(L, path)
    type(Layer), intent(inout) :: L
    character(len=*), intent(in) :: path
    integer(int64) :: bytes, n
    integer :: u
    type(Box), allocatable :: b(:)
    inquire(file=trim(path), size=bytes)
    n = bytes / 16
    allocate(b(max(n, 1_int64)))
    open(newunit=u, file=trim(path), access='stream', form='unformatted', status='old', action='read')
    read(u) b(1:n)
    close(u)
    call layer_set_boxes(L, b, n)
  end subroutine load

  subroutine save(L, name)
    type(Layer), intent(inout) :: L
    character(len=*), intent(in) :: name
    integer :: u
    call layer_to_host(L)
    open(newunit=u, file=trim(dir)//'/'//name, access='stream', form='unformatted', status='replace', &
         action='write')
    if (L%n_used > 0) write(u) L%layer_boxes(1:L%n_used)
    close(u)
  end subroutine save

  subroutine save_labels(L, name)
    type(Layer), intent(in) :: L
    character(len=*), intent(in) :: name
    integer :: u
    open(newunit=u, file=trim(dir)//'/'//name, access='stream', form='unformatted', status='replace', &
         action='write')
    if (L%n_used > 0) write(u) L%pnumtable%arr(1:L%n_used)
    close(u)
  end subroutine save_labels

end program check_layer_ops
