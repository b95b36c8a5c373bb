! modules.f90 -- the Fortran host side of the drop-in.  The reference's
! program uses four modules (glob_var.f90, compt_acc.f90, init_update.f90,
! in_output.f90) whose hot-path subroutines take no arguments and share module
! state.  The modules below keep those names and subroutine names; the state
! lives on the GPU behind the engine handle, and each subroutine forwards to
! the C-ABI.

module global_variables
  use iso_c_binding
  use md_api
  implicit none
  save
  type(md_params) :: prm                 ! every value of the parameter files
  type(c_ptr) :: engine = c_null_ptr
  integer :: NN, Nq, Npe, Nta, Ntl, Nml, Ngl, N_anchor
  integer :: StepNum0, StepNum, DeltaStep1, DeltaStep2, DeltaStep3, multistep
  integer :: restart_or_continue, step = 0
  real*8 :: Lx, Ly, Lz, dt
  integer :: SizeHist = 500
  real*8, allocatable, dimension(:,:) :: pos, vel, acc
  real*8 :: dr_max(2)
  real*8 :: started, finished, total_time = 0

contains

  ! Stops with the engine's message, as the reference stops on a bond break
  ! (compt_acc.f90:879-887).
  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    character(kind=c_char) :: buf(512)
    character(len=512) :: msg
    integer :: k
    integer(c_int) :: rc2
    if (rc == MD_OK) return
    msg = ' '
    rc2 = md_last_error(engine, buf, 512_c_int32_t)
    do k = 1, 512
      if (buf(k) == c_null_char) exit
      msg(k:k) = buf(k)
    end do
    if (rc == MD_EBOND) then
      write(*,*) 'The bond is break off!'
    end if
    write(*,*) trim(what), ': status', rc, ' ', trim(msg)
    stop 1
  end subroutine check

  subroutine fetch_state
    call check(md_download_state(engine, pos, vel, acc, dr_max), 'download')
  end subroutine fetch_state

end module global_variables


module compute_acceleration
  use global_variables
  implicit none
  save

contains

  ! initialize_force_parameters (compt_acc.f90:85-135): the engine builds
  ! bond, anchor and charge lists, B-spline tables and the influence
  ! function for the mesh and per-axis spline orders of force_data.txt and
  ! the reference's automatic mesh rule (compt_acc.f90:375-430).
  subroutine initialize_force_parameters
    character(len=16) :: env
    integer :: lenv, st
    prm%precision = MD_FP64
    call get_environment_variable('MD_PRECISION', env, lenv, st)
    if (st == 0 .and. env(1:4) == 'fp32') prm%precision = MD_FP32
    if (prm%seed == 0) prm%seed = 20261203_c_int64_t
    call check(md_create(engine, prm), 'md_create')
  end subroutine initialize_force_parameters

  ! lj_verlet_list + real_verlet_list (compt_acc.f90:185-314, 491-626)
  subroutine lj_verlet_list
    call check(md_build_lists(engine), 'lj_verlet_list')
  end subroutine lj_verlet_list

  subroutine real_verlet_list
    ! built together with the LJ list
  end subroutine real_verlet_list

  ! update_verlet_list (compt_acc.f90:699-726)
  subroutine update_verlet_list
    call check(md_update_verlet_list(engine), 'update_verlet_list')
  end subroutine update_verlet_list

  ! Compute_Force (compt_acc.f90:815-862)
  subroutine compute_force
    call check(md_compute_force(engine, int(step, c_int32_t), MD_NOISE_PHILOX, c_null_ptr), &
               'compute_force')
  end subroutine compute_force

  ! error_analysis (compt_acc.f90:729-812): the engine compares its Coulomb
  ! force with a standard-Ewald reference (tol = 5) on the device.  The
  ! messages and the line appended to rms_force.txt follow the reference;
  ! the two time columns hold the wall time of the whole device check and 0.
  subroutine error_analysis
    real(c_double) :: out(8)
    integer(c_int64_t) :: np1, np2
    integer(8) :: t0, t1, rate
    call system_clock(t0, rate)
    call check(md_error_analysis(engine, out, c_null_ptr), 'error_analysis')
    call system_clock(t1)
    write(*,*) 'Standard Ewald error in real space:', out(2)
    write(*,*) 'Standard Ewald error in fourier spqce', out(3)
    call check(md_download_list(engine, 0_c_int32_t, c_null_ptr, 0_c_int64_t, np1), 'lists')
    call check(md_download_list(engine, 1_c_int32_t, c_null_ptr, 0_c_int64_t, np2), 'lists')
    write(*,*) 'Rms force of coulomb force is:', out(1)
    open(60, position='append', file='./data/rms_force.txt')
      write(60, '(6F17.6)') step*1., out(1), np1*1., np2*1., dble(t1 - t0) / dble(rate), 0.d0
    close(60)
  end subroutine error_analysis

end module compute_acceleration


module initialize_update
  use global_variables
  use compute_acceleration
  implicit none

contains

  ! Initialize_position (init_update.f90:6-66), seeded.
  subroutine initialize_position
    call check(md_initialize_position(prm, prm%seed, pos), 'initialize_position')
    vel = 0
    acc = 0
    call check(md_upload_state(engine, pos, vel, acc), 'upload')
    call lj_verlet_list
    write(*,*) 'Inititalize position is finished!'
  end subroutine initialize_position

  ! initialize_velocity (init_update.f90:462-495)
  subroutine initialize_velocity
    call check(md_initialize_velocity(prm, prm%seed + 1_c_int64_t, vel), 'initialize_velocity')
    call check(md_upload_state(engine, pos, vel, acc), 'upload')
    write(*,*) 'initialize velocity finished'
  end subroutine initialize_velocity

  ! new_position (init_update.f90:555-607), with period_condition_pos inside
  subroutine new_position
    call check(md_new_position(engine, int(step, c_int32_t), MD_NOISE_PHILOX, c_null_ptr), &
               'new_position')
  end subroutine new_position

  ! rescale_velocity (init_update.f90:498-552)
  subroutine rescale_velocity
    call check(md_rescale_velocity(engine), 'rescale_velocity')
  end subroutine rescale_velocity

  subroutine period_condition_pos
    ! applied on the device inside new_position
  end subroutine period_condition_pos

  ! Steps step0..step0+n-1 of the loop body (new_position, pre-heat rescale,
  ! update_verlet_list) without leaving the device.
  subroutine run_steps(step0, n, preheat)
    integer, intent(in) :: step0, n, preheat
    if (n <= 0) return
    call check(md_run(engine, int(step0, c_int32_t), int(n, c_int32_t), int(preheat, c_int32_t)), &
               'run')
  end subroutine run_steps

end module initialize_update


module input_output
  use global_variables
  implicit none

contains

  ! initialize_parameters (in_output.f90:145-168): system_data.txt,
  ! force_data.txt and q_arm.txt, derived quantities, restart flag.
  subroutine initialize_parameters
    logical :: alive
    call check(md_read_params('./system_data.txt'//c_null_char, './force_data.txt'//c_null_char, &
                              './q_arm.txt'//c_null_char, prm), 'read parameters')
    call check(md_derive_params(prm), 'derive')
    NN = prm%NN; Nq = prm%Nq; Npe = prm%Npe; Nta = prm%Nta; Ntl = prm%Ntl
    Nml = prm%Nml; Ngl = prm%Ngl; N_anchor = prm%N_anchor
    StepNum0 = prm%StepNum0; StepNum = prm%StepNum
    DeltaStep1 = prm%DeltaStep1; DeltaStep2 = prm%DeltaStep2; DeltaStep3 = prm%DeltaStep3
    multistep = prm%multistep
    Lx = prm%Lx; Ly = prm%Ly; Lz = prm%Lz; dt = prm%dt
    allocate(pos(NN,4), vel(NN,3), acc(NN,3))
    inquire(file='start_time.txt', exist=alive)
    restart_or_continue = 0
    if (alive) then
      open(11, file='./start_time.txt')
        read(11,*) restart_or_continue
      close(11)
    end if
    write(*,*) 'NN =', NN, ' Nq =', Nq, ' Lx =', Lx
  end subroutine initialize_parameters

  ! continue_read_data (in_output.f90:605-618): positions, velocities, and
  ! the step after the last checkpoint.
  subroutine continue_read_data(l)
    integer, intent(out) :: l
    integer :: i
    real*8 :: t
    open(20, file='./data/pos1.txt')
      read(20,*) ((pos(i,l), l=1,4), i=1,NN)
    close(20)
    open(20, file='./data/vel1.txt')
      read(20,*) ((vel(i,l), l=1,3), i=1,NN)
    close(20)
    open(11, file='./start_time.txt')
      read(11,*) i
      read(11,*) l
      read(11,*) t
    close(11)
    l = l + 1
    total_time = t
    acc = 0
    call check(md_upload_state(engine, pos, vel, acc), 'upload')
  end subroutine continue_read_data

  ! Brush height: mean z of the polymer beads (in_output.f90:1017-1020).
  ! height (in_output.f90:1017-1079): h_avg, hl_max, hl_end, hl_avg,
  ! hs_max, hs_end, hs_branch, hs_avg, computed on the device.
  subroutine height(h)
    real*8, intent(out) :: h(8)
    call check(md_heights(engine, h), 'height')
  end subroutine height

  ! write_height's height.txt line (in_output.f90:1342-1344, format 9F17.6)
  subroutine write_height(j, h)
    integer, intent(in) :: j
    real*8, intent(in) :: h(8)
    open(36, position='append', file='./data/height.txt')
      write(36,'(9F17.6)') j*1., h
    close(36)
  end subroutine write_height

  ! z-density of the polymer beads, SizeHist bins of Lz/SizeHist, ceiling
  ! binning (in_output.f90:1878-1886); written as one profile per call.
  subroutine histogram
    integer :: i, k
    integer, allocatable :: phi(:)
    allocate(phi(SizeHist))
    phi = 0
    call fetch_state
    do i = 1, Npe
      k = ceiling(pos(i,3) / (Lz / SizeHist))
      if (k <= 0 .or. k > SizeHist) cycle
      phi(k) = phi(k) + 1
    end do
    open(37, position='append', file='./data/phi_tot.txt')
      do k = 1, SizeHist
        write(37,'(2F17.6)') k * Lz / SizeHist, phi(k) * 1.
      end do
    close(37)
    deallocate(phi)
  end subroutine histogram

  ! write_pos, write_vel, write_acc (in_output.f90:2428-2452, 2482-2507,
  ! 2551-2576): the current state as text, 4F17.6 / 3F17.6
  subroutine write_pos
    integer :: i
    call fetch_state
    open(30, file='./data/pos.txt')
      do i = 1, NN
        write(30,'(4F17.6)') pos(i,1), pos(i,2), pos(i,3), pos(i,4)
      end do
    close(30)
  end subroutine write_pos

  subroutine write_vel
    integer :: i
    call fetch_state
    open(32, file='./data/vel.txt')
      do i = 1, NN
        write(32,'(3F17.6)') vel(i,1), vel(i,2), vel(i,3)
      end do
    close(32)
  end subroutine write_vel

  subroutine write_acc
    integer :: i
    call fetch_state
    open(37, file='./data/acc.txt')
      do i = 1, NN
        write(37,'(3F17.6)') acc(i,1), acc(i,2), acc(i,3)
      end do
    close(37)
  end subroutine write_acc

  ! write_pos1 / write_vel1 (in_output.f90:2473-2478, 2529-2546): text
  ! checkpoint and the restart flag file.
  subroutine write_checkpoint(j)
    integer, intent(in) :: j
    integer :: i
    call fetch_state
    open(32, file='./data/pos1.txt')
      do i = 1, NN
        write(32,'(4F17.6)') pos(i,1), pos(i,2), pos(i,3), pos(i,4)
      end do
    close(32)
    open(33, file='./data/vel1.txt')
      do i = 1, NN
        write(33,'(3F17.6)') vel(i,1), vel(i,2), vel(i,3)
      end do
    close(33)
    call write_state
    call cpu_time(finished)
    open(34, file='./start_time.txt')
      write(34,*) 1
      write(34,*) j
      write(34,*) finished - started + total_time
    close(34)
  end subroutine write_checkpoint

  ! The engine's complete state beside the text checkpoint (data/state.bin,
  ! raw bytes): with MD_EXACT_RESUME=1 a restart loads it and continues bit
  ! for bit instead of the reference's text restart.
  subroutine write_state
    integer(c_int64_t) :: n
    integer(c_int8_t), allocatable :: buf(:)
    call check(md_state_size(engine, n), 'state size')
    allocate(buf(n))
    call check(md_state_save(engine, buf, n), 'state save')
    open(38, file='./data/state.bin', access='stream', form='unformatted', status='replace')
      write(38) buf
    close(38)
    deallocate(buf)
  end subroutine write_state

  ! Exact resume from data/state.bin when MD_EXACT_RESUME=1 and the file
  ! exists; .false. otherwise.
  logical function read_state()
    character(len=8) :: flag
    integer :: st
    integer(c_int64_t) :: n
    integer(c_int8_t), allocatable :: buf(:)
    logical :: alive
    read_state = .false.
    call get_environment_variable('MD_EXACT_RESUME', flag, status=st)
    if (st /= 0 .or. flag(1:1) /= '1') return
    inquire(file='./data/state.bin', exist=alive, size=n)
    if (.not. alive .or. n <= 0) return
    allocate(buf(n))
    open(38, file='./data/state.bin', access='stream', form='unformatted', status='old')
      read(38) buf
    close(38)
    call check(md_state_load(engine, buf, n), 'state load')
    deallocate(buf)
    read_state = .true.
  end function read_state

end module input_output
