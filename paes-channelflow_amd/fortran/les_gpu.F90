! les_gpu.F90 -- Fortran driver of the MI355X channel-flow LES.
!
! Drop-in for the reference driver program les (les.F90): it reads the same
! nlist.input namelists (RESTARTING, TIMING, GRIDSIZE, PARAMS; les.F90:30-34,
! defaults of default.f90), the same restart .field file (les.F90:69-83,
! read_write.f90:1-11), runs the same time loop and RK3 stage loop
! (les.F90:203-285) and writes the same outputs: the per-iteration wall clock,
! the mean-profile table, parameters.dat, <case><time>.mean
! (les.F90:287-323, 389-408) and <case><n>.field (les.F90:331-351).
! The tendency, pressure and update of every stage run on one GPU through the
! C-ABI of include/channelflow.h.
!
! One optional namelist group is added:
!   &GPU precision=0|1 (0: FP64, 1: FP32 fields with FP64 pressure),
!        device, alk (<= 0: dsl), seed, amp,
!        closure=0|1 (1: learned SGS closure with TKE fallback), weights='file',
!        closure_grid=0|1 (0: network on the 3/2 grid, 1: on the nnx x nny grid) /
! seed/amp select the seeded initial condition used when restart=.false.
! (the reference's initvel draws from the compiler's random_number).
! Restart files may hold REAL(4) records (the shipped reference) or REAL(8)
! records (a promoted build); the kind is detected from the first record.
! Field files are written with REAL(8) records.
module channelflow_c
  use iso_c_binding
  implicit none
  integer, parameter :: CF_PREC_FP64 = 0, CF_PREC_FP32 = 1

  type, bind(C) :: cf_params
     integer(c_int32_t) :: nnx, nny, nnz
     real(c_double) :: xl, yl, zl, z0, dt
     integer(c_int32_t) :: imodel, isurface, precision, device
     real(c_double) :: alk, ugal, vgal
  end type cf_params

  interface
     integer(c_int) function cf_create(p, ctx) bind(C, name="cf_create")
       import :: c_int, c_ptr, cf_params
       type(cf_params), intent(in) :: p
       type(c_ptr), intent(out) :: ctx
     end function cf_create
     integer(c_int) function cf_destroy(ctx) bind(C, name="cf_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function cf_destroy
     integer(c_int) function cf_set_state(ctx, u, v, w, e) bind(C, name="cf_set_state")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: u(*), v(*), w(*), e(*)
     end function cf_set_state
     integer(c_int) function cf_seeded_init(ctx, seed, amp) bind(C, name="cf_seeded_init")
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int64_t), value :: seed
       real(c_double), value :: amp
     end function cf_seeded_init
     integer(c_int) function cf_get_state(ctx, u, v, w, e, p) bind(C, name="cf_get_state")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: u(*), v(*), w(*), e(*), p(*)
     end function cf_get_state
     integer(c_int) function cf_stage(ctx, istage, mnout) bind(C, name="cf_stage")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: istage, mnout
     end function cf_stage
     integer(c_int) function cf_sync(ctx) bind(C, name="cf_sync")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function cf_sync
     integer(c_int) function cf_get_scalars(ctx, s) bind(C, name="cf_get_scalars")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: s(4)
     end function cf_get_scalars
     integer(c_int) function cf_compmn(ctx, meanbuf) bind(C, name="cf_compmn")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: meanbuf(*)
     end function cf_compmn
     type(c_ptr) function cf_last_error(ctx) bind(C, name="cf_last_error")
       import :: c_ptr
       type(c_ptr), value :: ctx
     end function cf_last_error
     ! learned SGS closure (no counterpart in the reference)
     integer(c_int) function cf_mlp_create(device, m) bind(C, name="cf_mlp_create")
       import :: c_int, c_int32_t, c_ptr
       integer(c_int32_t), value :: device
       type(c_ptr), intent(out) :: m
     end function cf_mlp_create
     integer(c_int) function cf_mlp_load(m, path) bind(C, name="cf_mlp_load")
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: m
       character(kind=c_char), intent(in) :: path(*)
     end function cf_mlp_load
     integer(c_int) function cf_mlp_destroy(m) bind(C, name="cf_mlp_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: m
     end function cf_mlp_destroy
     integer(c_int) function cf_set_closure(ctx, m) bind(C, name="cf_set_closure")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, m
     end function cf_set_closure
     integer(c_int) function cf_set_closure_grid(ctx, grid) bind(C, name="cf_set_closure_grid")
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: grid
     end function cf_set_closure_grid
     integer(c_int) function cf_get_closure_counts(ctx, counts) bind(C, name="cf_get_closure_counts")
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int64_t), intent(out) :: counts(2)
     end function cf_get_closure_counts
  end interface

contains

  ! lesabort (abort.F90:1-17): print the library's message and stop
  subroutine check(rc, ctx, what)
    integer(c_int), intent(in) :: rc
    type(c_ptr), intent(in) :: ctx
    character(*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    type(c_ptr) :: m
    integer :: n
    if (rc == 0) return
    m = cf_last_error(ctx)
    n = 0
    if (c_associated(m)) then
       call c_f_pointer(m, msg, [512])
       do while (n < 512)
          if (msg(n+1) == c_null_char) exit
          n = n + 1
       end do
       write(*,*) what, " failed (", rc, "): ", transfer(msg(1:n), repeat(" ", n))
    else
       write(*,*) what, " failed (", rc, ")"
    end if
    stop 1
  end subroutine check

end module channelflow_c


program les_gpu
  use iso_c_binding
  use channelflow_c
  implicit none

  ! namelist variables (les.F90:30-34, parameters.f90, modules.f90)
  logical :: restart, isbgr, spectout, spectav
  character(100) :: restart_file, case
  integer :: timax, savfrq, imodel, isurface, nnx, nny, nnz
  real(8) :: meanfrq, cfl, dt, cb, skc, xl, yl, zl, z0
  integer :: precision, device, closure, closure_grid
  integer(8) :: seed
  real(8) :: alk, amp
  character(256) :: weights
  namelist /RESTARTING/ restart, restart_file
  namelist /TIMING/ case, timax, savfrq, meanfrq, isbgr, cfl, dt, &
       cb, skc, spectout, spectav, imodel, isurface
  namelist /GRIDSIZE/ nnx, nny, nnz
  namelist /PARAMS/ xl, yl, zl, z0
  namelist /GPU/ precision, device, alk, seed, amp, closure, weights, closure_grid

  type(cf_params) :: p
  type(c_ptr) :: ctx, net
  integer(c_int64_t) :: counts(2)
  real(8), allocatable :: u(:,:,:), v(:,:,:), w(:,:,:), e(:,:,:), pr(:,:,:), meanbuf(:,:)
  real(8) :: time, time_start, utau, divgls, fcor, ugtop, ugbot, vgtop, vgbot, s4(4)
  real(8) :: seconds, wallclock0, wallclock1, zm, zt
  integer :: it, istage, iz, ios, meancount, fieldcount, rk, mnout_i
  integer :: time_array(8)
  logical :: mnout, fdout, exists
  character(25) :: filename, meanfile

  ! defaults (default.f90:6-30)
  restart = .false.; restart_file = "restart.bin"; spectout = .false.; spectav = .true.
  case = "les"; timax = 100; savfrq = 100; meanfrq = 100; isbgr = .true.
  nnx = 32; nny = 32; nnz = 32; xl = 5000.0d0; yl = 5000.0d0; zl = 2000.0d0; z0 = 0.16d0
  cfl = 0.3d0; dt = 0.01d0; cb = 0; skc = 0; imodel = 3; isurface = 1
  precision = CF_PREC_FP64; device = 0; alk = 0.0d0; seed = 10972_8; amp = 0.1d0
  closure = 0; weights = "sgs_mlp.cfmlp"; net = c_null_ptr; closure_grid = 0
  divgls = 0.0d0; fcor = 0.0001d0; ugtop = 15.0d0; ugbot = 15.0d0; vgtop = 0.0d0; vgbot = 0.0d0

  write(*,*)
  write(*,*) "NCAR/PENN STATE LARGE-EDDY SIMULATION (MI355X)"
  write(*,*)
  call date_and_time(values=time_array)
  wallclock0 = time_array(5)*3600.d0 + time_array(6)*60.d0 + time_array(7) + 0.001d0*time_array(8)
  seconds = 0.0d0

  inquire(file="nlist.input", exist=exists)
  if (.not. exists) then
     write(*,*) "Stopping; can't find nlist.input"
     stop 1
  end if
  open(10, file="nlist.input", action="read")
  read(10, nml=RESTARTING)
  read(10, nml=TIMING)
  if (restart) then
     inquire(file=restart_file, exist=exists)
     if (.not. exists) then
        write(*,*) "Stopping; can't find restart file "//trim(restart_file)
        stop 1
     end if
     write(*,*) "Reading parameters from restart file"
     rk = record_kind(restart_file)
     open(50, file=restart_file, form="unformatted", action="read")
     call read_header(rk, time_start, nnx, nny, nnz, xl, yl, zl, dt, z0, utau)
  else
     write(*,*) "Reading parameters from namelist file"
     read(10, nml=GRIDSIZE)
     read(10, nml=PARAMS)
     time_start = 0.0d0
  end if
  rewind(10)
  read(10, nml=GPU, iostat=ios)
  close(10)
  if (.not. isbgr) then
     write(*,*) "Stopping; isbgr=.false. is not supported (the SGS TKE is always carried)"
     stop 1
  end if

  p%nnx = nnx; p%nny = nny; p%nnz = nnz
  p%xl = xl; p%yl = yl; p%zl = zl; p%z0 = z0; p%dt = dt
  p%imodel = imodel; p%isurface = isurface; p%precision = precision; p%device = device
  p%alk = alk; p%ugal = 0.0d0; p%vgal = 0.0d0
  call check(cf_create(p, ctx), c_null_ptr, "cf_create")
  if (closure == 1) then
     call check(cf_mlp_create(int(device, c_int32_t), net), c_null_ptr, "cf_mlp_create")
     if (cf_mlp_load(net, trim(weights)//c_null_char) /= 0) then
        write(*,*) "Stopping; cannot load the closure network ", trim(weights)
        stop 1
     end if
     call check(cf_set_closure(ctx, net), ctx, "cf_set_closure")
     call check(cf_set_closure_grid(ctx, int(closure_grid, c_int32_t)), ctx, "cf_set_closure_grid")
     write(*,*) "Learned SGS closure from ", trim(weights), " (TKE fallback)"
  end if

  allocate(u(nnx,nny,nnz), v(nnx,nny,nnz), w(nnx,nny,nnz), e(nnx,nny,nnz), pr(nnx,nny,nnz))
  allocate(meanbuf(19,nnz))
  if (restart) then
     do iz = 1, nnz                                  ! read_field (read_write.f90:1-11)
        call read_level(rk, u(:,:,iz), v(:,:,iz), w(:,:,iz), e(:,:,iz), pr(:,:,iz))
     end do
     close(50)
     call check(cf_set_state(ctx, u, v, w, e), ctx, "cf_set_state")
  else
     call check(cf_seeded_init(ctx, seed, amp), ctx, "cf_seeded_init")
  end if

  it = 0
  time = 0.0d0
  meancount = 1
  fieldcount = 1
  do while (time < dble(timax))                       ! les.F90:203
     time = time + dt
     call date_and_time(values=time_array)
     wallclock1 = time_array(5)*3600.d0 + time_array(6)*60.d0 + time_array(7) + 0.001d0*time_array(8)
     if (wallclock1 >= wallclock0) then
        seconds = seconds + (wallclock1 - wallclock0)
     else
        seconds = seconds + (86400.0d0 - wallclock0) + wallclock1
     end if
     wallclock0 = wallclock1
     write(*,*) "Iteration ", it+1, "; model time = ", time+time_start, "; delta t = ", dt
     write(*,*) "Wall Clock Time = ", seconds, "seconds"
     write(*,*)

     if (time >= dble(meancount)*meanfrq) then       ! les.F90:238-245
        meancount = meancount + 1
        mnout = .true.
     else if (it == 0) then
        mnout = .true.
     else
        mnout = .false.
     end if
     if (time >= dble(fieldcount*savfrq)) then       ! les.F90:247-252
        fieldcount = fieldcount + 1
        fdout = .true.
     else
        fdout = .false.
     end if

     mnout_i = merge(1, 0, mnout)
     do istage = 1, 3                                ! les.F90:254-285
        call check(cf_stage(ctx, istage, mnout_i), ctx, "cf_stage")
     end do
     call check(cf_sync(ctx), ctx, "cf_sync")

     if (mnout .or. it == 0) then                    ! les.F90:287-323
        call check(cf_compmn(ctx, meanbuf), ctx, "cf_compmn")
        call check(cf_get_scalars(ctx, s4), ctx, "cf_get_scalars")
        write(filename,'(F9.2)') time + time_start
        meanfile = trim(case)//trim(adjustl(filename))//".mean"
        write(*,'(a4,3a9,3a11)') "IZ", "Uxym ", "Vxym ", "pple ", "engle", "divz  ", "engsbz "
        do iz = 1, nnz
           write(*,'(I4,3f9.3,3e11.3)') iz, meanbuf(4,iz), meanbuf(5,iz), meanbuf(13,iz), &
                meanbuf(1,iz), meanbuf(3,iz), meanbuf(2,iz)
        end do
        open(500, file='parameters.dat', position='append')      ! les.F90:389-408
        if (it == 0) write(500,'(5a12)') '#Time', 'utau', 'utau_s', 'dudt', 'dvdt'
        write(500,'(f12.3, 4f12.6 )') time+time_start, s4(1), s4(2), s4(3), s4(4)
        close(500)
        if (mnout .and. meancount >= 1) then
           open(unit=35, file=meanfile)
           write(35,*) nnz, time+time_start
           write(35,*) s4(1), z0
           write(35,*) ((iz-0.5d0)*(zl/nnz), iz=1,nnz), (iz*(zl/nnz), iz=1,nnz)
           write(35,*) meanbuf(1,:), meanbuf(2,:)
           write(35,*) meanbuf(4,:), meanbuf(5,:), meanbuf(6,:)
           write(35,*) meanbuf(7,:), meanbuf(8,:), meanbuf(9,:), meanbuf(10,:), meanbuf(11,:), meanbuf(12,:)
           write(35,*) meanbuf(13,:)
           write(35,*) meanbuf(14,:), meanbuf(15,:), meanbuf(17,:), meanbuf(16,:), meanbuf(18,:), meanbuf(19,:)
           close(35)
        end if
     end if

     if (fdout) then                                 ! les.F90:331-351
        call check(cf_get_state(ctx, u, v, w, e, pr), ctx, "cf_get_state")
        call check(cf_get_scalars(ctx, s4), ctx, "cf_get_scalars")
        write(filename,'(I10)') (fieldcount-1)*savfrq + int(time_start)
        filename = trim(case)//trim(adjustl(filename))//".field"
        open(unit=14, file=filename, form="unformatted")
        write(14) time+time_start, nnx, nny, nnz, xl, yl, zl           ! read_write.f90:22-24
        write(14) dt, z0, s4(1)
        write(14) divgls, fcor, ugtop, ugbot, vgtop, vgbot
        do iz = 1, nnz
           write(14) u(:,:,iz), v(:,:,iz), w(:,:,iz), e(:,:,iz), pr(:,:,iz)
        end do
        close(14)
     end if
     it = it + 1
  end do

  if (closure == 1) then
     call check(cf_get_closure_counts(ctx, counts), ctx, "cf_get_closure_counts")
     write(*,'(a,i0,a,i0,a)') " Learned closure used at ", counts(1), " of ", counts(2), &
          " cell evaluations since the last output"
     call check(cf_set_closure(ctx, c_null_ptr), ctx, "cf_set_closure")
     ios = cf_mlp_destroy(net)
  end if
  call check(cf_destroy(ctx), c_null_ptr, "cf_destroy")
  call date_and_time(values=time_array)
  wallclock1 = time_array(5)*3600.d0 + time_array(6)*60.d0 + time_array(7) + 0.001d0*time_array(8)
  if (wallclock1 >= wallclock0) then
     seconds = seconds + (wallclock1 - wallclock0)
  else
     seconds = seconds + (86400.0d0 - wallclock0) + wallclock1
  end if
  write(*,*) "Job took", seconds, "seconds to execute"
  write(*,*)

contains

  ! 4 for REAL(4) records, 8 for REAL(8): the first record holds
  ! (time, nnx, nny, nnz, xl, yl, zl) = 4 reals + 3 default integers.
  integer function record_kind(fname)
    character(*), intent(in) :: fname
    integer(4) :: marker
    open(51, file=fname, access="stream", form="unformatted", action="read")
    read(51) marker
    close(51)
    if (marker == 4*4 + 3*4) then
       record_kind = 4
    else
       record_kind = 8
    end if
  end function record_kind

  subroutine read_header(rk, t, nx, ny, nz, lx, ly, lz, dtv, z0v, ut)
    integer, intent(in) :: rk
    real(8), intent(out) :: t, lx, ly, lz, dtv, z0v, ut
    integer, intent(out) :: nx, ny, nz
    real(4) :: s(6)
    if (rk == 4) then
       read(50) s(1), nx, ny, nz, s(2), s(3), s(4)
       t = s(1); lx = s(2); ly = s(3); lz = s(4)
       read(50) s(1), s(2), s(3)
       dtv = s(1); z0v = s(2); ut = s(3)
       read(50) s(1:6)
    else
       read(50) t, nx, ny, nz, lx, ly, lz
       read(50) dtv, z0v, ut
       read(50) divgls, fcor, ugtop, ugbot, vgtop, vgbot
    end if
  end subroutine read_header

  subroutine read_level(rk, a, b, c, d, q)
    integer, intent(in) :: rk
    real(8), intent(out) :: a(:,:), b(:,:), c(:,:), d(:,:), q(:,:)
    real(4), allocatable :: t(:,:,:)
    if (rk == 4) then
       allocate(t(size(a,1), size(a,2), 5))
       read(50) t(:,:,1), t(:,:,2), t(:,:,3), t(:,:,4), t(:,:,5)
       a = t(:,:,1); b = t(:,:,2); c = t(:,:,3); d = t(:,:,4); q = t(:,:,5)
    else
       read(50) a, b, c, d, q
    end if
  end subroutine read_level

end program les_gpu
