! main.f90 -- the reference's main program flow (main.f90:1-115) over the
! GPU engine: fresh start or restart, pre-heat with velocity rescaling every
! 20 steps below step 300000, production, analysis and checkpoint cadences.
! Between two analysis steps the loop body (new_position, rescale,
! update_verlet_list) runs on the device without host round trips.
program main
  use global_variables
  use initialize_update
  use compute_acceleration
  use input_output
  implicit none
  integer :: i, nxt, last
  real*8 :: h(8)

  call cpu_time(started)
  call initialize_parameters
  call initialize_force_parameters

  if (restart_or_continue == 0) then
    i = 1
    call initialize_position
    call initialize_velocity
    call write_checkpoint(0)
    if (prm%qq /= 0) call error_analysis
    step = 0
    call compute_force
    call write_acc
  else
    call continue_read_data(i)
    if (read_state()) then  ! exact resume: the saved state, lists and forces
      write(*,*) 'exact resume from data/state.bin at step', i
      step = i
    else
      call lj_verlet_list
      if (prm%qq /= 0) call error_analysis
      call rescale_velocity
      ! the reference's step counter is still 0 here (main.f90:58 runs before
      ! any loop sets it), so this force includes a fresh Coulomb evaluation
      ! that the following steps hold until their first multiple of multistep
      step = 0
      call compute_force
      call write_pos
      call write_vel
      call write_acc
    end if
  end if

  !##############preheating##############!
  if (i <= StepNum0) then
    step = i
    do while (step <= StepNum0)
      nxt = min(next_mult(step, DeltaStep1), next_mult(step, DeltaStep3), StepNum0)
      call run_steps(step, nxt - step + 1, 1)
      step = nxt
      if (mod(step, DeltaStep1) == 0) then
        call height(h)
        call write_height(step, h)
      end if
      if (mod(step, DeltaStep3) == 0) then
        if (prm%qq /= 0) call error_analysis
        call write_checkpoint(step)
      end if
      step = step + 1
    end do
    i = step
  end if

  !################running###############!
  last = StepNum0 + StepNum
  step = i
  do while (step <= last)
    nxt = min(next_mult(step, DeltaStep1), next_mult(step, DeltaStep2), next_mult(step, DeltaStep3), last)
    call run_steps(step, nxt - step + 1, 0)
    step = nxt
    if (mod(step, DeltaStep1) == 0) then
      call height(h)
      call write_height(step, h)
    end if
    if (mod(step, DeltaStep2) == 0) call histogram
    if (mod(step, DeltaStep3) == 0) call write_checkpoint(step)
    step = step + 1
  end do

  call fetch_state
  call cpu_time(finished)
  write(*,*) 'steps:', last, ' height:', sum(pos(1:Npe,3)) / Npe
  write(*,*) 'Finished!'
  call md_destroy(engine)

contains

  integer function next_mult(s, d)
    integer, intent(in) :: s, d
    next_mult = ((s + d - 1) / d) * d
  end function next_mult

end program main
