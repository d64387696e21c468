! vl_control.f90 — an `exec run` interpreter over vl_layer_ops.
!
! Statement forms follow the reference's control language for the operator
! path (src/controlparser.f90:341-593):
!   exec run LHS = RHS1 <op> RHS2       <op> one of + @ * ~ ! ^ %   (:409-476)
!   exec run LHS = RHS1 NAME RHS2|nothing [args]                    (:478-585)
!       NAME: OR2 AND2 NOT FRAMENOT XOR COPY EXTENT ASSERT_ZERO
!             AND (one input: the points two or more boxes cover)
!             GRID [rows cols overlap]  (the reference's fixed 10 10 10)
!             GROW e n w s   SIZE d    (lengths in microns, times the
!             precision and truncated to DBU as in :566-577)
!   decl design D = FILE.mag      an input design (:257-271): every layer
!   exec modedesign D = FILE.mag  section of the .mag file becomes layer D:NAME,
!                                 from its rect lines or the KLBIN / SNAP /
!                                 DSNAP file it names (src/magparser.f90:
!                                 108-290); each layer is then numbered and,
!                                 when its boxes overlap, healed (:358-418)
!   decl output O = FILE.mag      an output design (:240-256): its sections
!                                 name the files flush writes (KLBIN for .bin,
!                                 SNAP for .snap, src/magparser.f90:510-564)
!   decl memory M = nothing       a design of intermediate layers (:272-289)
!   exec init DB:NAME             an empty layer (:310-340)
!   var temp_folder | temporary_layers | abort_on_assert_zero | abort_on_xor |
!       debug_verbosity VALUE     (:180-235): a layer created by exec run or
!                                 exec init in a declared design is backed by
!                                 temp_folder//DB//'_MTL_'//NAME;
!                                 temporary_layers 2 spills memory designs to
!                                 those files at flush and reads a missing
!                                 layer file as an empty layer
!   program NAME, systeminfo, info  accepted (messages only)
!   flush                         run the deferred operators, then write the
!                                 output designs' non-empty layers to their
!                                 files and release them (:613-626)
!   delete NAME                   release a layer, or every layer of a design
!                                 when NAME ends in ':' (:627-642); a released
!                                 layer with a backing file is read again when
!                                 next used (lazy restore, :440-441); delete D
!                                 of a declared design drops its layers
! plus the statements a stand-alone driver adds:
!   precision N                   DBU per micron (default 1000)
!   layer NAME klbin FILE         raw little-endian int32 boxes (src/kldata.f90)
!   layer NAME synth CFG SEED K N seeded synthetic layer (include/vl_synth.h)
!   dump NAME FILE                write the layer's boxes as KLBIN
!   area NAME                     print the exact union area
!   resident on|off               keep results on the device between statements
!   schedule immediate|batch|graph
!                                 immediate: every exec run runs when read;
!                                 batch (default): consecutive exec run
!                                 statements are recorded into one program
!                                 (independent chains run concurrently,
!                                 operator groups over one layer fuse; the
!                                 reference notes that a graph could be built,
!                                 :308-309) and run at the next statement that
!                                 needs their results (dump, area, layer,
!                                 delete, flush, an operator programs do not
!                                 take, the end of the file); graph: the same,
!                                 with the program captured as a HIP graph
! Layer names are taken whole ("d1:poly"); "nothing" and "<db>:nothing" name
! no layer.  '#' starts a comment.  Unknown statements are reported and
! skipped, as the reference does (:658-659).
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
     character(len=:), allocatable :: file   ! backing file (lazy restore, flush)
     logical :: file_is_klbin = .false.       ! read with `layer NAME klbin`
     logical :: released = .false.
     integer :: db = 0                        ! declared design, 0 for none
  end type NamedLayer

  ! Declared designs (src/design.f90:86-88 directions)
  integer, parameter :: DIR_INPUT = 0, DIR_OUTPUT = 1, DIR_MEMORY = 2, DIR_DELETED = -1
  type :: DeclaredDesign
     character(len=:), allocatable :: name
     character(len=:), allocatable :: file
     integer :: direction = DIR_INPUT
  end type DeclaredDesign

  type(NamedLayer), allocatable, target :: table(:)
  integer :: nlayers = 0
  type(DeclaredDesign), allocatable :: designs(:)
  integer :: ndesigns = 0
  character(len=:), allocatable :: temp_folder
  integer(int64) :: used_precision = 1000
  integer :: assert_failures = 0
  integer :: statements_run = 0

  ! deferred operators (schedule batch / graph)
  integer :: schedule_mode = 1   ! 0 immediate, 1 batch, 2 graph
  type(c_ptr), save :: prog = c_null_ptr
  integer :: npend = 0
  integer, allocatable :: pend_out(:)           ! layers the pending program writes
  integer :: npend_out = 0
  integer, allocatable :: pend_assert(:, :)     ! (layer, line) of pending ASSERT_ZEROs
  integer :: npend_assert = 0
  integer(c_int), parameter :: OP_XOR = 0, OP_OR = 1, OP_AND = 2, OP_NOT = 3, OP_MERGE = 4, OP_SIZE = 5, &
       OP_GROW = 6

  interface
     function c_program_create(e, p) bind(C, name="vl_program_create") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: e
       type(c_ptr) :: p
       integer(c_int) :: rc
     end function c_program_create
     function c_program_destroy(p) bind(C, name="vl_program_destroy") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: p
       integer(c_int) :: rc
     end function c_program_destroy
     function c_program_add(p, op, a, b, c, args) bind(C, name="vl_program_add") result(rc)
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: p, a, b, c
       integer(c_int), value :: op
       integer(c_int32_t) :: args(4)
       integer(c_int) :: rc
     end function c_program_add
     function c_program_run(p) bind(C, name="vl_program_run") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: p
       integer(c_int) :: rc
     end function c_program_run
     function c_program_set_graph(p, enable) bind(C, name="vl_program_set_graph") result(rc)
       import :: c_int, c_ptr
       type(c_ptr), value :: p
       integer(c_int), value :: enable
       integer(c_int) :: rc
     end function c_program_set_graph
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
          if (allocated(table(i)%file)) call move_alloc(table(i)%file, grown(i)%file)
          grown(i)%file_is_klbin = table(i)%file_is_klbin
          grown(i)%released = table(i)%released
          grown(i)%db = table(i)%db
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
    if (allocated(src%pnumtable%arr)) call move_alloc(src%pnumtable%arr, dst%pnumtable%arr)
    dst%area = src%area
    dst%perimeter = src%perimeter
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
          if (table(i)%released) then
             if (.not. allocated(table(i)%file)) then
                write(error_unit, '(A,A,A,I0)') 'ERROR: layer was deleted: ', trim(name), ' on line ', line_no
                error stop "ERROR: deleted layer used"
             end if
             if (table(i)%file_is_klbin) then       ! lazy restore
                call load_klbin(table(i)%L, table(i)%file)
             else
                call RestoreSnapToLayer(table(i)%L, table(i)%file)
             end if
             table(i)%released = .false.
          end if
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

  ! ---- deferred operators -----------------------------------------------------

  subroutine note_output(il)
    integer, intent(in) :: il
    integer, allocatable :: g(:)
    integer :: k
    do k = 1, npend_out
       if (pend_out(k) == il) return
    end do
    if (.not. allocated(pend_out)) allocate(pend_out(64))
    if (npend_out == size(pend_out)) then
       allocate(g(2 * size(pend_out)))
       g(1:npend_out) = pend_out(1:npend_out)
       call move_alloc(g, pend_out)
    end if
    npend_out = npend_out + 1
    pend_out(npend_out) = il
  end subroutine note_output

  ! Records op(table(ia), table(ib)) -> table(il) into the pending program.
  subroutine defer(op, ia, ib, il, args)
    integer(c_int), intent(in) :: op
    integer, intent(in) :: ia, ib, il
    integer(c_int32_t), intent(in) :: args(4)
    type(c_ptr) :: ha, hb, hc
    integer(c_int32_t) :: a(4)
    if (.not. c_associated(prog)) call vl_check(c_program_create(vl_engine_handle(), prog), "vl_program_create")
    ha = layer_input_dev(table(ia)%L)
    hb = c_null_ptr
    if (ib > 0) hb = layer_input_dev(table(ib)%L)
    hc = layer_output_dev(table(il)%L)
    a = args
    call vl_check(c_program_add(prog, op, ha, hb, hc, a), "vl_program_add")
    call note_output(il)
    npend = npend + 1
  end subroutine defer

  subroutine defer_assert(i1, line_no)
    integer, intent(in) :: i1, line_no
    integer, allocatable :: g(:, :)
    if (.not. allocated(pend_assert)) allocate(pend_assert(2, 64))
    if (npend_assert == size(pend_assert, 2)) then
       allocate(g(2, 2 * size(pend_assert, 2)))
       g(:, 1:npend_assert) = pend_assert(:, 1:npend_assert)
       call move_alloc(g, pend_assert)
    end if
    npend_assert = npend_assert + 1
    pend_assert(:, npend_assert) = [i1, line_no]
  end subroutine defer_assert

  subroutine report_assert(i1, line_no)
    integer, intent(in) :: i1, line_no
    if (LayerAssertZero(table(i1)%L) > 0) then
       write(output_unit, '(A,I5,A)') '************ ASSERT_ZERO fail on layer. Line: ', line_no, &
            ' **********************'
       assert_failures = assert_failures + 1
    end if
  end subroutine report_assert

  ! Runs the pending program (if any), then the deferred ASSERT_ZERO checks
  ! in statement order.
  subroutine flush_pending()
    integer :: k
    if (npend > 0) then
       if (schedule_mode == 2) call vl_check(c_program_set_graph(prog, 1_c_int), "vl_program_set_graph")
       call vl_check(c_program_run(prog), "vl_program_run")
       do k = 1, npend_out
          call layer_output_done(table(pend_out(k))%L)
       end do
    end if
    do k = 1, npend_assert
       call report_assert(pend_assert(1, k), pend_assert(2, k))
    end do
    if (c_associated(prog)) call vl_check(c_program_destroy(prog), "vl_program_destroy")
    prog = c_null_ptr
    npend = 0
    npend_out = 0
    npend_assert = 0
  end subroutine flush_pending

  ! delete NAME / delete DB: (every layer whose name starts with "DB:").
  subroutine delete_layers(name)
    character(len=*), intent(in) :: name
    integer :: i, n
    logical :: all_of_db
    n = len_trim(name)
    all_of_db = name(n:n) == ':'
    do i = 1, nlayers
       if (all_of_db) then
          if (len(table(i)%name) < n) cycle
          if (table(i)%name(1:n) /= name(1:n)) cycle
       else
          if (table(i)%name /= trim(name)) cycle
       end if
       call layer_release(table(i)%L)
       if (allocated(table(i)%L%layer_boxes)) deallocate(table(i)%L%layer_boxes)
       table(i)%L%n_used = 0
       table(i)%L%n_alloc = 0
       table(i)%released = .true.
    end do
  end subroutine delete_layers

  ! ---- designs ----------------------------------------------------------------

  integer function design_index(name)
    character(len=*), intent(in) :: name
    integer :: k
    design_index = 0
    do k = 1, ndesigns
       if (designs(k)%direction == DIR_DELETED) cycle
       if (designs(k)%name == trim(name)) then
          design_index = k
          return
       end if
    end do
  end function design_index

  ! Declared design of a layer name "DB:NAME" (0 when DB is not declared).
  integer function design_of(name)
    character(len=*), intent(in) :: name
    integer :: p
    design_of = 0
    p = index(name, ':')
    if (p > 1) design_of = design_index(name(1:p-1))
  end function design_of

  logical function layer_exists(name)
    character(len=*), intent(in) :: name
    integer :: i
    layer_exists = .false.
    do i = 1, nlayers
       if (table(i)%name == trim(name)) then
          layer_exists = .true.
          return
       end if
    end do
  end function layer_exists

  integer function declare_design(name, file, direction)
    character(len=*), intent(in) :: name, file
    integer, intent(in) :: direction
    type(DeclaredDesign), allocatable :: g(:)
    integer :: k
    if (design_index(name) > 0) then
       write(error_unit, '(A,A)') 'ERROR: Duplicate db handle seen: ', trim(name)
       error stop "ERROR: duplicate design"
    end if
    if (.not. allocated(designs)) allocate(designs(8))
    if (ndesigns == size(designs)) then
       allocate(g(2 * size(designs)))
       do k = 1, ndesigns
          call move_alloc(designs(k)%name, g(k)%name)
          if (allocated(designs(k)%file)) call move_alloc(designs(k)%file, g(k)%file)
          g(k)%direction = designs(k)%direction
       end do
       call move_alloc(g, designs)
    end if
    ndesigns = ndesigns + 1
    designs(ndesigns)%name = trim(name)
    designs(ndesigns)%file = trim(file)
    designs(ndesigns)%direction = direction
    declare_design = ndesigns
  end function declare_design

  ! A layer created in declared design k is backed by
  ! temp_folder//DB//'_MTL_'//NAME (src/controlparser.f90:322-330,363-373).
  subroutine back_new_layer(i, k)
    integer, intent(in) :: i, k
    integer :: p
    character(len=:), allocatable :: lname
    p = index(table(i)%name, ':')
    lname = table(i)%name(p+1:)
    table(i)%db = k
    table(i)%file_is_klbin = .false.
    if (lname == 'nothing') then
       table(i)%file = 'MTL_NOTHING'
    else
       table(i)%file = temp_folder//designs(k)%name//'_MTL_'//lname
    end if
  end subroutine back_new_layer

  ! The layer an exec run writes (:341-384): created when new (backed as
  ! above in a declared design); a released layer is replaced by the result.
  integer function lhs_index(name)
    character(len=*), intent(in) :: name
    integer :: k
    logical :: fresh
    fresh = .not. layer_exists(name)
    lhs_index = layer_index(name)
    table(lhs_index)%released = .false.
    k = design_of(name)
    if (fresh .and. k > 0) call back_new_layer(lhs_index, k)
  end function lhs_index

  ! exec init DB:NAME (:310-340): an empty layer of a declared design.
  subroutine exec_init(name, line_no)
    character(len=*), intent(in) :: name
    integer, intent(in) :: line_no
    integer :: i, k
    k = design_of(name)
    if (k == 0) then
       write(error_unit, '(A,A,A,I0)') 'DB: ', trim(name), ' not found. Syntax ERROR: line: ', line_no
       error stop "Syntax ERROR"
    end if
    i = layer_index(name)
    call back_new_layer(i, k)
    call ClearLayer(table(i)%L)
    table(i)%released = .false.
  end subroutine exec_init

  ! addBoxToLayer (src/magparser.f90:493-507): appends one valid box.
  subroutine add_box(L, b)
    type(Layer), intent(inout) :: L
    type(Box), intent(in) :: b
    type(Box), allocatable :: g(:)
    if (.not. (b%x1 < b%x2 .and. b%y1 < b%y2)) then
       write(error_unit, '(A,4I12)') 'Box ', b%x1, b%y1, b%x2, b%y2
       error stop "ERROR: box in valid"
    end if
    if (.not. L%host_valid) call layer_to_host(L)
    if (.not. allocated(L%layer_boxes)) allocate(L%layer_boxes(4))
    if (L%n_used >= size(L%layer_boxes, kind=int64)) then
       allocate(g(2 * size(L%layer_boxes, kind=int64)))
       g(1:L%n_used) = L%layer_boxes(1:L%n_used)
       call move_alloc(g, L%layer_boxes)
    end if
    L%n_used = L%n_used + 1
    L%layer_boxes(L%n_used) = b
    L%n_alloc = size(L%layer_boxes, kind=int64)
    call layer_boxes_changed(L)
  end subroutine add_box

  ! Reads design k's .mag file (parseMagicLayoutFile, src/magparser.f90:
  ! 108-290): each section "<< NAME >>" (but labels and end) is layer
  ! DB:NAME; rect lines add boxes to it; KLBIN / SNAP / DSNAP lines name its
  ! backing file, read now for an input design.  "<< end >>" ends the file;
  ! magscale, rlabel and other lines are skipped.  An input design's layers
  ! are then numbered and, when their boxes overlap, healed (:358-418).
  subroutine load_mag(k)
    integer, intent(in) :: k
    character(len=4096) :: line
    character(len=256) :: dummy
    character(len=:), allocatable :: section
    integer :: u, ios, cur, i, n
    integer(int32) :: x1, y1, x2, y2
    open(newunit=u, file=designs(k)%file, status='old', action='read', iostat=ios)
    if (ios /= 0) then
       write(error_unit, '(A,A)') 'ERROR: cannot open design file ', designs(k)%file
       error stop "ERROR: design file"
    end if
    cur = 0
    do
       read(u, '(A)', iostat=ios) line
       if (ios /= 0) exit
       if (len_trim(line) == 0 .or. line(1:1) == '#') cycle
       if (line(1:8) == 'magscale') cycle
       if (line(1:9) == '<< end >>') exit
       if (line(1:2) == '<<') then
          n = len_trim(line)
          section = trim(adjustl(line(4:max(3, n - 2))))
          if (section == 'labels' .or. section == 'end') cycle
          cur = layer_index(designs(k)%name//':'//section)
          table(cur)%db = k
          table(cur)%released = .false.
          cycle
       end if
       if (cur == 0) cycle
       if (line(1:4) == 'rect') then
          read(line, *, iostat=ios) dummy, x1, y1, x2, y2
          if (ios /= 0) then
             write(error_unit, '(A,A)') 'ERROR: bad rect line in ', designs(k)%file
             error stop "ERROR: design file"
          end if
          call add_box(table(cur)%L, Box(x1, y1, x2, y2))
       else if (line(1:5) == 'KLBIN' .or. line(1:5) == 'DSNAP') then
          call attach_file(cur, k, trim(adjustl(line(7:))))
       else if (line(1:4) == 'SNAP') then
          call attach_file(cur, k, trim(adjustl(line(6:))))
       end if
    end do
    close(u)
    if (designs(k)%direction /= DIR_INPUT) return
    do i = 1, nlayers
       if (table(i)%db == k .and. table(i)%L%n_used > 0) call PNumHealLayer(table(i)%L)
    end do
  end subroutine load_mag

  subroutine attach_file(i, k, f)
    integer, intent(in) :: i, k
    character(len=*), intent(in) :: f
    table(i)%file = f
    table(i)%file_is_klbin = .false.
    if (designs(k)%direction == DIR_INPUT) call RestoreSnapToLayer(table(i)%L, f)
  end subroutine attach_file

  ! flush (:613-626, writeMagicLayoutFile src/magparser.f90:510-564): every
  ! non-empty layer of an output design goes to its file (SNAP when the name
  ! ends in .snap, KLBIN when it ends in .bin, anything else is an error),
  ! and with temporary_layers 2 every non-empty layer of a memory design to
  ! its SNAP file; each written layer is released (read again when used).
  subroutine flush_designs()
    integer :: i, k, n
    character(len=:), allocatable :: f
    do i = 1, nlayers
       k = table(i)%db
       if (k == 0) cycle
       if (designs(k)%direction /= DIR_OUTPUT .and. &
            .not. (designs(k)%direction == DIR_MEMORY .and. temporary_layers == 2)) cycle
       if (table(i)%released .or. table(i)%L%n_used == 0) cycle
       if (.not. allocated(table(i)%file)) then
          write(error_unit, '(A,A)') 'Layer not assigned backing store: ', table(i)%name
          error stop "ERROR: layer backing store name not allocated"
       end if
       f = table(i)%file
       n = len(f)
       if (designs(k)%direction == DIR_MEMORY) then
          call SaveLayerToSnap(table(i)%L, f, int(K_COMPRESSION_METHOD_TO_USE))
       else if (n >= 5 .and. f(max(1, n - 4):n) == '.snap') then
          call SaveLayerToSnap(table(i)%L, f, int(K_COMPRESSION_METHOD_TO_USE))
       else if (n >= 4 .and. f(max(1, n - 3):n) == '.bin') then
          call layer_to_host(table(i)%L)
          call WriteKLBin(f, table(i)%L%layer_boxes, table(i)%L%n_used)
       else
          write(error_unit, '(A,A)') 'ERROR: File format not supported: ', f
          error stop "DBOUT only supports KLBIN, edit the output MAG file to fix"
       end if
       call ClearLayer(table(i)%L)
       table(i)%released = .true.
    end do
  end subroutine flush_designs

  ! delete D of a declared design (:627-642, DeleteDesign): its layers are
  ! dropped; using one afterwards is an error.
  subroutine delete_design(k)
    integer, intent(in) :: k
    integer :: i
    do i = 1, nlayers
       if (table(i)%db /= k) cycle
       call ClearLayer(table(i)%L)
       if (allocated(table(i)%file)) deallocate(table(i)%file)
       table(i)%released = .true.
       table(i)%db = 0
    end do
    designs(k)%direction = DIR_DELETED
  end subroutine delete_design

  ! Splits s into blank-separated words.
  subroutine split_words(s, w, nw)
    character(len=*), intent(in) :: s
    character(len=256), intent(out) :: w(:)
    integer, intent(out) :: nw
    integer :: i, j, n
    w = ' '
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
    integer :: il, i1, i2, k, rows, cols, ovl
    integer(int32) :: iv(4)
    integer(int64) :: d
    integer(c_int) :: code
    logical :: later
    ! w: exec run LHS = RHS1 OP RHS2 [args]
    if (nw < 7 .or. w(4) /= '=') then
       write(error_unit, '(A,I0)') 'SYNTAX ERROR on line: ', line_no
       error stop "SYNTAX ERROR on exec run"
    end if
    later = schedule_mode /= 0
    opw = w(6)
    i1 = existing(w(5), line_no)
    if (len_trim(opw) == 1 .and. scan(opw, '+@*%^~!') == 1) then
       if (opw(1:1) == '%') return
       i2 = existing(w(7), line_no)
       il = lhs_index(w(3))
       select case (opw(1:1))
       case ('+')
          code = OP_OR
       case ('*')
          code = OP_AND
       case ('^')
          code = OP_XOR
       case default   ! '@', '~', '!': the same point set A \ B
          code = OP_NOT
       end select
       if (later) then
          call defer(code, i1, i2, il, [0, 0, 0, 0])
       else
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
       end if
       statements_run = statements_run + 1
       return
    end if
    select case (trim(opw))
    case ('ASSERT_ZERO')
       if (.not. is_nothing(w(7))) error stop "ASSERT_ZERO must use nothing as second layer"
       if (later) then
          call defer_assert(i1, line_no)
       else
          call report_assert(i1, line_no)
       end if
    case ('OR2', 'AND2', 'NOT', 'FRAMENOT', 'XOR')
       i2 = existing(w(7), line_no)
       il = lhs_index(w(3))
       if (later) then
          select case (trim(opw))
          case ('OR2')
             code = OP_OR
          case ('AND2')
             code = OP_AND
          case ('XOR')
             code = OP_XOR
          case default
             code = OP_NOT
          end select
          call defer(code, i1, i2, il, [0, 0, 0, 0])
       else
          select case (trim(opw))
          case ('OR2')
             call CalculateOR(table(i1)%L, table(i2)%L, table(il)%L)
          case ('AND2')
             call CalculateFrameAND(table(i1)%L, table(i2)%L, table(il)%L)
          case ('NOT')
             call CalculateNOT(table(i1)%L, table(i2)%L, table(il)%L)
          case ('FRAMENOT')
             call CalculateFrameNOT(table(i1)%L, table(i2)%L, table(il)%L)
          case ('XOR')
             call CalculateXOR(table(i1)%L, table(i2)%L, table(il)%L)
          end select
       end if
    case ('COPY')
       il = lhs_index(w(3))
       if (later) then
          call defer(OP_MERGE, i1, 0, il, [0, 0, 0, 0])
       else
          call CopyLayer(table(il)%L, table(i1)%L)
       end if
    case ('EXTENT', 'AND', 'GRID')
       ! operators a program does not take: run what is pending, then this one
       call flush_pending()
       il = lhs_index(w(3))
       select case (trim(opw))
       case ('EXTENT')
          if (.not. is_nothing(w(7))) error stop "EXTENT must use nothing as second layer"
          call CreateEXTENT(table(i1)%L, table(il)%L)
       case ('AND')
          call CalculateSingleLayerAND(table(i1)%L, table(il)%L)
       case ('GRID')
          rows = 10
          cols = 10
          ovl = 10
          if (nw >= 10) then
             rows = nint(to_real(w(8), line_no))
             cols = nint(to_real(w(9), line_no))
             ovl = nint(to_real(w(10), line_no))
          end if
          call CreateGrid(table(i1)%L, table(il)%L, rows, cols, ovl)
       end select
    case ('GROW')
       if (.not. is_nothing(w(7))) error stop "GROW must use nothing as second layer"
       if (nw < 11) error stop "SYNTAX ERROR: GROW needs four lengths"
       do k = 1, 4
          iv(k) = int(real(used_precision, real64) * to_real(w(7 + k), line_no), int32)
       end do
       il = lhs_index(w(3))
       if (later) then
          call defer(OP_GROW, i1, 0, il, iv)
       else
          call CalculateGROWLayer(table(i1)%L, table(il)%L, iv)
       end if
    case ('SIZE')
       if (.not. is_nothing(w(7))) error stop "SIZE must use nothing as second layer"
       if (nw < 8) error stop "SYNTAX ERROR: SIZE needs a length"
       d = int(to_real(w(8), line_no) * real(used_precision, real64), int64)
       il = lhs_index(w(3))
       if (later) then
          if (d < -huge(0_int32) - 1_int64 .or. d > huge(0_int32)) error stop "SIZE amount out of int32 range"
          call defer(OP_SIZE, i1, 0, il, [int(d, int32), 0, 0, 0])
       else
          call CalculateBoostOperation(table(i1)%L, table(i1)%L, table(il)%L, K_BOOST_CONTROL_SIZE, d)
       end if
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
    if (.not. allocated(temp_folder)) temp_folder = ''
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
       ! every statement other than exec run needs the results so far
       if (trim(w(1)) /= 'exec' .or. trim(w(2)) /= 'run') call flush_pending()
       select case (trim(w(1)))
       case ('exec')
          select case (trim(w(2)))
          case ('run')
             call exec_run(w, nw, line_no)
          case ('init')
             if (nw < 3) error stop "SYNTAX ERROR: exec init DB:NAME"
             call exec_init(w(3), line_no)
          case ('modedesign')
             if (nw < 5 .or. w(4) /= '=') error stop "SYNTAX ERROR: exec modedesign D = FILE"
             call load_mag(declare_design(w(3), w(5), DIR_INPUT))
          case default
             write(error_unit, '(A,I0)') 'Syntax ERROR: line: ', line_no
             error stop "SYNTAX ERROR: exec run | init | modedesign"
          end select
       case ('decl')
          if (nw < 5 .or. w(4) /= '=') then
             write(error_unit, '(A,I0)') 'Syntax ERROR: line: ', line_no
             error stop "Syntax ERROR."
          end if
          select case (trim(w(2)))
          case ('design')
             call load_mag(declare_design(w(3), w(5), DIR_INPUT))
          case ('output')
             call load_mag(declare_design(w(3), w(5), DIR_OUTPUT))
          case ('memory')
             if (trim(w(5)) /= 'nothing') error stop "MEMORY DB must have backing store nothing"
             i = declare_design(w(3), 'nothing', DIR_MEMORY)
          case default
             write(error_unit, '(A,I0)') 'Syntax ERROR: line: ', line_no
             error stop "Syntax ERROR."
          end select
       case ('var')
          select case (trim(w(2)))
          case ('temp_folder')
             temp_folder = trim(w(3))
          case ('temporary_layers')
             temporary_layers = nint(to_real(w(3), line_no))
          case default   ! abort_on_assert_zero, abort_on_xor, debug_verbosity: accepted
             continue
          end select
       case ('program')
          write(output_unit, '(A,A)') '--- Starting Program: ', trim(w(2))
       case ('systeminfo', 'info')
          continue
       case ('schedule')
          select case (trim(w(2)))
          case ('immediate')
             schedule_mode = 0
          case ('batch')
             schedule_mode = 1
          case ('graph')
             schedule_mode = 2
          case default
             error stop "SYNTAX ERROR: schedule immediate|batch|graph"
          end select
       case ('flush')
          call flush_designs()
       case ('delete')
          if (nw < 2) error stop "SYNTAX ERROR: delete NAME"
          if (design_index(w(2)) > 0) then
             call delete_design(design_index(w(2)))
          else
             call delete_layers(w(2))
          end if
       case ('end')
          exit
       case ('precision')
          used_precision = nint(to_real(w(2), line_no), int64)
       case ('layer')
          if (nw < 4) error stop "SYNTAX ERROR: layer NAME klbin FILE | synth CFG SEED K N"
          i = layer_index(w(2))
          if (w(3) == 'klbin') then
             call load_klbin(table(i)%L, w(4))
             table(i)%file = trim(w(4))
             table(i)%file_is_klbin = .true.
             table(i)%released = .false.
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
          write(output_unit, '(A,A,A,I0)') 'Unknown command: ', trim(w(1)), ' on line ', line_no
       end select
    end do
    close(u)
    call flush_pending()
    call vl_fortran_sync()
  end subroutine run_control_file

end module vl_control
