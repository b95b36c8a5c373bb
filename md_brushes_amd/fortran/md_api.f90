! md_api.f90 -- iso_c_binding interface to the engine's C-ABI
! (include/md_engine.h).  md_params mirrors the C struct field for field.
module md_api
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: MD_OK = 0, MD_EINVAL = 1, MD_EBOND = 2, MD_ECELL = 3
  integer(c_int), parameter :: MD_ENONFINITE = 4, MD_EOVERFLOW = 5, MD_EHIP = 6
  integer(c_int), parameter :: MD_FP64 = 0, MD_FP32 = 1
  integer(c_int), parameter :: MD_NOISE_NONE = 0, MD_NOISE_ETA = 1, MD_NOISE_PHILOX = 2

  type, bind(C) :: md_params
    integer(c_int32_t) :: uniform_or_random, arm, Nma, Nga, Nml, Ngl, man_l, man_s
    integer(c_int32_t) :: StepNum0, StepNum, DeltaStep1, DeltaStep2, DeltaStep3, multistep
    real(c_double) :: Lz, Z_empty, sigmag, Beta, qq, qqi, ion_ratio, R_bond, dt
    real(c_double) :: epsilon, sigma, sigma_s, rcl, rvl, rsk, R0_2, kFENE
    integer(c_int32_t) :: ordr(3), pad0
    real(c_double) :: lb, xi, EF, tol, tau_rf
    integer(c_int32_t) :: charged_arm(16)
    integer(c_int32_t) :: ovr_ng(3), pad1
    real(c_double) :: ovr_alpha, ovr_rcc
    integer(c_int32_t) :: Nta, Ntl, Npe, N_anchor, Nq_PE, Nq_salt_ions, Nq, NN, N_bond
    integer(c_int32_t) :: real_verlet, ng(3), pad2
    real(c_double) :: Lx, Ly, Zfac, sigma_sp, rcl_s, rcl_sp, rvl_eff, alpha, rcc, rvc, gdim(3)
    integer(c_int32_t) :: precision, pad3
    integer(c_int64_t) :: seed
    real(c_double) :: ovr_skin, ovr_cell_scale
    integer(c_int32_t) :: ovr_tile(3), pad4
  end type md_params

  interface
    integer(c_int) function md_read_params(sysf, forcef, qarmf, p) bind(C, name="md_read_params")
      import :: c_int, c_char, md_params
      character(kind=c_char), dimension(*), intent(in) :: sysf, forcef, qarmf
      type(md_params), intent(inout) :: p
    end function
    integer(c_int) function md_derive_params(p) bind(C, name="md_derive_params")
      import :: c_int, md_params
      type(md_params), intent(inout) :: p
    end function
    integer(c_int) function md_initialize_position(p, seed, pos) bind(C, name="md_initialize_position")
      import :: c_int, c_int64_t, c_double, md_params
      type(md_params), intent(in) :: p
      integer(c_int64_t), value :: seed
      real(c_double), intent(out) :: pos(*)
    end function
    integer(c_int) function md_initialize_velocity(p, seed, vel) bind(C, name="md_initialize_velocity")
      import :: c_int, c_int64_t, c_double, md_params
      type(md_params), intent(in) :: p
      integer(c_int64_t), value :: seed
      real(c_double), intent(out) :: vel(*)
    end function
    integer(c_int) function md_create(e, p) bind(C, name="md_create")
      import :: c_int, c_ptr, md_params
      type(c_ptr), intent(out) :: e
      type(md_params), intent(in) :: p
    end function
    integer(c_int) function md_upload_state(e, pos, vel, acc) bind(C, name="md_upload_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: e
      real(c_double), intent(in) :: pos(*), vel(*), acc(*)
    end function
    integer(c_int) function md_download_state(e, pos, vel, acc, dr) bind(C, name="md_download_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: e
      real(c_double), intent(out) :: pos(*), vel(*), acc(*), dr(2)
    end function
    integer(c_int) function md_build_lists(e) bind(C, name="md_build_lists")
      import :: c_int, c_ptr
      type(c_ptr), value :: e
    end function
    integer(c_int) function md_compute_force(e, step, noise, eta) bind(C, name="md_compute_force")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: e, eta
      integer(c_int32_t), value :: step, noise
    end function
    integer(c_int) function md_new_position(e, step, noise, eta) bind(C, name="md_new_position")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: e, eta
      integer(c_int32_t), value :: step, noise
    end function
    integer(c_int) function md_update_verlet_list(e) bind(C, name="md_update_verlet_list")
      import :: c_int, c_ptr
      type(c_ptr), value :: e
    end function
    integer(c_int) function md_rescale_velocity(e) bind(C, name="md_rescale_velocity")
      import :: c_int, c_ptr
      type(c_ptr), value :: e
    end function
    integer(c_int) function md_run(e, step0, nsteps, preheat) bind(C, name="md_run")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: e
      integer(c_int32_t), value :: step0, nsteps, preheat
    end function
    integer(c_int) function md_state_size(e, n) bind(C, name="md_state_size")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: e
      integer(c_int64_t), intent(out) :: n
    end function
    integer(c_int) function md_state_save(e, buf, cap) bind(C, name="md_state_save")
      import :: c_int, c_ptr, c_int64_t, c_int8_t
      type(c_ptr), value :: e
      integer(c_int8_t), dimension(*), intent(out) :: buf
      integer(c_int64_t), value :: cap
    end function
    integer(c_int) function md_state_load(e, buf, n) bind(C, name="md_state_load")
      import :: c_int, c_ptr, c_int64_t, c_int8_t
      type(c_ptr), value :: e
      integer(c_int8_t), dimension(*), intent(in) :: buf
      integer(c_int64_t), value :: n
    end function
    integer(c_int) function md_heights(e, out) bind(C, name="md_heights")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: e
      real(c_double), dimension(8), intent(out) :: out
    end function
    integer(c_int) function md_error_analysis(e, out, acc_se) bind(C, name="md_error_analysis")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: e
      real(c_double), dimension(8), intent(out) :: out
      type(c_ptr), value :: acc_se
    end function
    integer(c_int) function md_download_list(e, which, pairs, cap, npair) &
        bind(C, name="md_download_list")
      import :: c_int, c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: e
      integer(c_int32_t), value :: which
      type(c_ptr), value :: pairs
      integer(c_int64_t), value :: cap
      integer(c_int64_t), intent(out) :: npair
    end function
    integer(c_int) function md_last_error(e, buf, n) bind(C, name="md_last_error")
      import :: c_int, c_ptr, c_char, c_int32_t
      type(c_ptr), value :: e
      character(kind=c_char), dimension(*), intent(out) :: buf
      integer(c_int32_t), value :: n
    end function
    subroutine md_destroy(e) bind(C, name="md_destroy")
      import :: c_ptr
      type(c_ptr), value :: e
    end subroutine
  end interface
end module md_api
