! vl_control.f90 — an `exec run` interpreter over vl_layer_ops.
!
! Statement forms follow the reference's control language for the operator
! path (src/controlparser.f90:341-593):
!   exec run LHS = RHS1 <op> RHS2       <op> one of + @ * ~ ! ^ %   (:409-476)
!   exec run LHS = RHS1 NAME RHS2|nothing [args]                    (:478-585)
!       NAME: OR2 AND2 NOT FRAMENOT XOR COPY EXTENT ASSERT_ZERO
!             GROW e n w s   SIZE d    (lengths in microns, times the
!             precision and truncated to DBU as in :566-577)
! plus the few statements a stand-alone driver needs in place of the
! reference's design loading (modedesign / .mag, which are out of this path):
!   precision N                   DBU per micron (default 1000)
!   layer NAME klbin FILE         raw little-endian int32 boxes (src/kldata.f90)
!   layer NAME synth CFG SEED K N seeded synthetic layer (include/vl_synth.h)
!   dump NAME FILE                write the layer's boxes as KLBIN
!   area NAME                     print the exact union area
!   resident on|off               keep results on the device between statements
! Layer names are taken whole ("d1:poly"); "nothing" and "<db>:nothing" name
! no layer.  '#' starts a comment.
module vl_control
  use iso_c_binding
  use iso_fortran_env, only: int32, int64, real64, output_unit, error_unit
  use vl_layer_ops
  implicit none
  private
  public :: run_control_file, assert_failures, statements_run

  type :: NamedLayer
     character(len=:), allocatable :: name
     type(Layer) :: L
  end type NamedLayer

  type(NamedLayer), allocatable, target :: table(:)
  integer :: nlayers = 0
  integer(int64) :: used_precision = 1000
  integer :: assert_failures = 0
  integer :: statements_run = 0

  interface
     function c_synth_layer(config, seed, layer, n, out) bind(C, name="vl_synth_layer") result(rc)
       import :: c_int, c_int64_t, c_ptr
       integer(c_int), value :: config, layer
       integer(c_int64_t), value :: seed, n
       type(c_ptr), value :: out
       integer(c_int) :: rc
     end function c_synth_layer
  end interface

contains

  logical function is_nothing(name)
    character(len=*), intent(in) :: name
    integer :: p
    p = index(name, ':', back=.true.)
    is_nothing = trim(name(p+1:)) == 'nothing'
  end function is_nothing

  ! Index of a layer by name, created empty when absent.
  integer function layer_index(name)
    character(len=*), intent(in) :: name
    type(NamedLayer), allocatable :: grown(:)
    integer :: i
    do i = 1, nlayers
       if (table(i)%name == trim(name)) then
          layer_index = i
          return
       end if
    end do
    if (.not. allocated(table)) allocate(table(16))
    if (nlayers == size(table)) then
       allocate(grown(2 * size(table)))
       do i = 1, nlayers
          call move_alloc(table(i)%name, grown(i)%name)
          call move_layer(table(i)%L, grown(i)%L)
       end do
       call move_alloc(grown, table)
    end if
    nlayers = nlayers + 1
    table(nlayers)%name = trim(name)
    layer_index = nlayers
  end function layer_index

  subroutine move_layer(src, dst)
    type(Layer), intent(inout) :: src, dst
    dst%n_used = src%n_used
    dst%n_alloc = src%n_alloc
    if (allocated(src%layer_boxes)) call move_alloc(src%layer_boxes, dst%layer_boxes)
    dst%layerState = src%layerState
    dst%dev = src%dev
    dst%dev_valid = src%dev_valid
    dst%host_valid = src%host_valid
    src%dev = c_null_ptr
  end subroutine move_layer

  integer function existing(name, line_no)
    character(len=*), intent(in) :: name
    integer, intent(in) :: line_no
    integer :: i
    do i = 1, nlayers
       if (table(i)%name == trim(name)) then
          existing = i
          return
       end if
    end do
    write(error_unit, '(A,A,A,I0)') 'ERROR: layer not found: ', trim(name), ' on line ', line_no
    error stop "ERROR: layer not found"
  end function existing

  subroutine load_klbin(L, path)
    type(Layer), intent(inout) :: L
    character(len=*), intent(in) :: path
    integer(int64) :: bytes, n
    integer :: u, ios
    type(Box), allocatable :: b(:)
    inquire(file=trim(path), size=bytes)
    if (bytes < 0) then
       write(error_unit, *) 'ERROR: KLBIN file not found: ', trim(path)
       error stop "ERROR: KLBIN file"
    end if
    n = bytes / 16
    allocate(b(max(n, 1_int64)))
    if (n > 0) then
       open(newunit=u, file=trim(path), access='stream', form='unformatted', status='old', action='read', iostat=ios)
       if (ios /= 0) error stop "ERROR: cannot open KLBIN file"
       read(u, iostat=ios) b(1:n)
       close(u)
       if (ios /= 0) error stop "ERROR: cannot read KLBIN file"
    end if
    call layer_set_boxes(L, b, n)
  end subroutine load_klbin

  subroutine write_klbin(L, path)
    type(Layer), intent(inout) :: L
    character(len=*), intent(in) :: path
    integer :: u, ios
    call layer_to_host(L)
    open(newunit=u, file=trim(path), access='stream', form='unformatted', status='replace', action='write', &
         iostat=ios)
    if (ios /= 0) error stop "ERROR: cannot open dump file"
    if (L%n_used > 0) write(u) L%layer_boxes(1:L%n_used)
    close(u)
  end subroutine write_klbin

  subroutine load_synth(L, cfg, seed, k, n)
    type(Layer), intent(inout) :: L
    integer, intent(in) :: cfg, k
    integer(int64), intent(in) :: seed, n
    type(Box), allocatable, target :: b(:)
    allocate(b(max(n, 1_int64)))
    if (c_synth_layer(int(cfg, c_int), seed, int(k, c_int), n, c_loc(b(1))) /= 0) &
         error stop "ERROR: synthetic layer generation failed"
    call layer_set_boxes(L, b, n)
  end subroutine load_synth

  ! EXTENT (src/controlparser.f90:538-543): the layer's bounding box.
  subroutine extent(src, dst)
    type(Layer), intent(inout) :: src, dst
    type(Box) :: m(1)
    integer(int64) :: i
    call layer_to_host(src)
    if (src%n_used == 0) then
       call layer_set_boxes(dst, m, 0_int64)
       return
    end if
    m(1) = src%layer_boxes(1)
    do i = 2, src%n_used
       m(1)%x1 = min(m(1)%x1, src%layer_boxes(i)%x1)
       m(1)%y1 = min(m(1)%y1, src%layer_boxes(i)%y1)
       m(1)%x2 = max(m(1)%x2, src%layer_boxes(i)%x2)
       m(1)%y2 = max(m(1)%y2, src%layer_boxes(i)%y2)
    end do
    call layer_set_boxes(dst, m, 1_int64)
  end subroutine extent

  ! Splits s into blank-separated words.
  subroutine split_words(s, w, nw)
    character(len=*), intent(in) :: s
    character(len=256), intent(out) :: w(:)
    integer, intent(out) :: nw
    integer :: i, j, n
    nw = 0
    n = len_trim(s)
    i = 1
    do while (i <= n .and. nw < size(w))
       do while (i <= n)
          if (s(i:i) /= ' ' .and. s(i:i) /= char(9)) exit
          i = i + 1
       end do
       if (i > n) exit
       j = i
       do while (j <= n)
          if (s(j:j) == ' ' .or. s(j:j) == char(9)) exit
          j = j + 1
       end do
       nw = nw + 1
       w(nw) = s(i:j-1)
       i = j
    end do
  end subroutine split_words

  real(real64) function to_real(word, line_no)
    character(len=*), intent(in) :: word
    integer, intent(in) :: line_no
    integer :: ios
    read(word, *, iostat=ios) to_real
    if (ios /= 0) then
       write(error_unit, '(A,A,A,I0)') 'SYNTAX ERROR: bad number ', trim(word), ' on line ', line_no
       error stop "SYNTAX ERROR"
    end if
  end function to_real

  subroutine exec_run(w, nw, line_no)
    character(len=256), intent(in) :: w(:)
    integer, intent(in) :: nw, line_no
    character(len=256) :: opw
    integer :: il, i1, i2, k
    integer(int32) :: iv(4)
    integer(int64) :: d
    ! w: exec run LHS = RHS1 OP RHS2 [args]
    if (nw < 7 .or. w(4) /= '=') then
       write(error_unit, '(A,I0)') 'SYNTAX ERROR on line: ', line_no
       error stop "SYNTAX ERROR on exec run"
    end if
    opw = w(6)
    i1 = existing(w(5), line_no)
    if (len_trim(opw) == 1 .and. scan(opw, '+@*%^~!') == 1) then
       if (opw(1:1) == '%') return
       i2 = existing(w(7), line_no)
       il = layer_index(w(3))
       select case (opw(1:1))
       case ('+')
          call CalculateBoostOperation(table(i1)%L, table(i2)%L, table(il)%L, K_BOOST_CONTROL_OR, 0_int64)
       case ('*')
          call CalculateAND(table(i1)%L, table(i2)%L, table(il)%L)
       case ('@')
          call CalculateNOT(table(i1)%L, table(i2)%L, table(il)%L)
       case ('~')
          call CalculateFrameNOT(table(i1)%L, table(i2)%L, table(il)%L)
       case ('!')
          call CalculateBoostOperation(table(i1)%L, table(i2)%L, table(il)%L, K_BOOST_CONTROL_NOT, 0_int64)
       case ('^')
          call CalculateBoostOperation(table(i1)%L, table(i2)%L, table(il)%L, K_BOOST_CONTROL_XOR, 0_int64)
       end select
       statements_run = statements_run + 1
       return
    end if
    select case (trim(opw))
    case ('ASSERT_ZERO')
       if (.not. is_nothing(w(7))) error stop "ASSERT_ZERO must use nothing as second layer"
       if (LayerAssertZero(table(i1)%L) > 0) then
          write(output_unit, '(A,I5,A)') '************ ASSERT_ZERO fail on layer. Line: ', line_no, &
               ' **********************'
          assert_failures = assert_failures + 1
       end if
    case ('OR2', 'AND2', 'NOT', 'FRAMENOT', 'XOR')
       i2 = existing(w(7), line_no)
       il = layer_index(w(3))
       select case (trim(opw))
       case ('OR2')
          call CalculateOR(table(i1)%L, table(i2)%L, table(il)%L)
       case ('AND2')
          call CalculateAND(table(i1)%L, table(i2)%L, table(il)%L)
       case ('NOT')
          call CalculateNOT(table(i1)%L, table(i2)%L, table(il)%L)
       case ('FRAMENOT')
          call CalculateFrameNOT(table(i1)%L, table(i2)%L, table(il)%L)
       case ('XOR')
          call CalculateXOR(table(i1)%L, table(i2)%L, table(il)%L)
       end select
    case ('COPY')
       il = layer_index(w(3))
       call CopyLayer(table(il)%L, table(i1)%L)
    case ('EXTENT')
       if (.not. is_nothing(w(7))) error stop "EXTENT must use nothing as second layer"
       il = layer_index(w(3))
       call extent(table(i1)%L, table(il)%L)
    case ('GROW')
       if (.not. is_nothing(w(7))) error stop "GROW must use nothing as second layer"
       if (nw < 11) error stop "SYNTAX ERROR: GROW needs four lengths"
       do k = 1, 4
          iv(k) = int(real(used_precision, real64) * to_real(w(7 + k), line_no), int32)
       end do
       il = layer_index(w(3))
       call CalculateGROWLayer(table(i1)%L, table(il)%L, iv)
    case ('SIZE')
       if (.not. is_nothing(w(7))) error stop "SIZE must use nothing as second layer"
       if (nw < 8) error stop "SYNTAX ERROR: SIZE needs a length"
       d = int(to_real(w(8), line_no) * real(used_precision, real64), int64)
       il = layer_index(w(3))
       call CalculateBoostOperation(table(i1)%L, table(i1)%L, table(il)%L, K_BOOST_CONTROL_SIZE, d)
    case default
       write(error_unit, '(A,A,A,I0)') 'UNKNOWN primary_operator: ', trim(opw), ' on line: ', line_no
       error stop "SYNTAX ERROR on TEXT based operator"
    end select
    statements_run = statements_run + 1
  end subroutine exec_run

  ! Runs every statement of a control file; returns the ASSERT_ZERO failures.
  subroutine run_control_file(path)
    character(len=*), intent(in) :: path
    character(len=4096) :: line
    character(len=256) :: w(16)
    integer :: u, ios, nw, line_no, p, i
    integer(int64) :: a
    open(newunit=u, file=trim(path), status='old', action='read', iostat=ios)
    if (ios /= 0) then
       write(error_unit, *) 'ERROR: cannot open control file ', trim(path)
       error stop "ERROR: control file"
    end if
    line_no = 0
    do
       read(u, '(A)', iostat=ios) line
       if (ios /= 0) exit
       line_no = line_no + 1
       p = index(line, '#')
       if (p > 0) line(p:) = ' '
       call split_words(line, w, nw)
       if (nw == 0) cycle
       select case (trim(w(1)))
       case ('exec')
          if (nw < 2 .or. w(2) /= 'run') error stop "SYNTAX ERROR: expected exec run"
          call exec_run(w, nw, line_no)
       case ('precision')
          used_precision = nint(to_real(w(2), line_no), int64)
       case ('layer')
          if (nw < 4) error stop "SYNTAX ERROR: layer NAME klbin FILE | synth CFG SEED K N"
          i = layer_index(w(2))
          if (w(3) == 'klbin') then
             call load_klbin(table(i)%L, w(4))
          else if (w(3) == 'synth' .and. nw >= 7) then
             call load_synth(table(i)%L, nint(to_real(w(4), line_no)), int(to_real(w(5), line_no), int64), &
                  nint(to_real(w(6), line_no)), int(to_real(w(7), line_no), int64))
          else
             error stop "SYNTAX ERROR: layer NAME klbin FILE | synth CFG SEED K N"
          end if
       case ('dump')
          call write_klbin(table(existing(w(2), line_no))%L, w(3))
       case ('area')
          a = LayerKleeArea(table(existing(w(2), line_no))%L)
          write(output_unit, '(A,A,A,I0)') 'AREA ', trim(w(2)), ' = ', a
       case ('resident')
          vl_eager_host = trim(w(2)) /= 'on'
       case default
          write(error_unit, '(A,A,A,I0)') 'SYNTAX ERROR: unknown statement ', trim(w(1)), ' on line ', line_no
          error stop "SYNTAX ERROR"
       end select
    end do
    close(u)
    call vl_fortran_sync()
  end subroutine run_control_file

end module vl_control
