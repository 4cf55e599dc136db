!> ISO_C_BINDING interface to the MI355X timestep library (include/lb3d_gpu.h).
!>
!> The derived type lb3d_params mirrors the C struct field for field; every
!> function returns 0 on success (LB3D_OK) and a nonzero code otherwise, with
!> the message from lb3d_gpu_error_message().
module lb3d_gpu_module
  use iso_c_binding
  implicit none
  private

  integer(c_int32_t), parameter, public :: LB3D_ABI_VERSION = 1
  integer(c_int32_t), parameter, public :: LB3D_OK = 0
  integer(c_int32_t), parameter, public :: LB3D_EINVAL = 1, LB3D_EHIP = 2, LB3D_ESTATE = 3, LB3D_ENOMEM = 4
  integer(c_int32_t), parameter, public :: LB3D_FP64 = 0, LB3D_FP32_DEV = 1
  integer(c_int32_t), parameter, public :: LB3D_BC_PERIODIC = 0, LB3D_BC_LEES_EDWARDS = 1

  type, bind(C), public :: lb3d_params
     integer(c_int32_t) :: nx, ny, nz
     integer(c_int32_t) :: ncomp
     integer(c_int32_t) :: bdist
     integer(c_int32_t) :: precision
     integer(c_int32_t) :: boundary
     integer(c_int32_t) :: restore
     real(c_double) :: amass(3)
     real(c_double) :: tau(3)
     real(c_double) :: tau_d, beta, d_0
     real(c_double) :: g_br, g_bs, g_ss
     real(c_double) :: shear_u
     real(c_double) :: shear_omega
     real(c_double) :: shear_sum
     real(c_double) :: g_accn(3)
     integer(c_int32_t) :: g_accn_min(3)
     integer(c_int32_t) :: g_accn_max(3)
     real(c_double) :: rho0(3)
     real(c_double) :: g_wr, g_wb, tau_wr, tau_wb
     real(c_double) :: rock_colour_r, rock_colour_b
     integer(c_int32_t) :: inv_fluid, inv_type
     real(c_double) :: inv_flux(3)
  end type lb3d_params

  !> lb3d_sanity: C [species][axis][xyz] arrays appear here as (xyz, axis, species)
  type, bind(C), public :: lb3d_sanity
     real(c_double) :: mass(3)
     real(c_double) :: momentum(3)
     real(c_double) :: vel_sum(3, 3)
     real(c_double) :: max_vel(3, 3)
     integer(c_int32_t) :: max_pos(3, 3, 3)
     integer(c_int32_t) :: pad0
     integer(c_int64_t) :: fluid_sites, nan_sites, negative_sites, fast_sites
  end type lb3d_sanity

  public :: lb3d_gpu_abi_version, lb3d_gpu_create, lb3d_gpu_destroy
  public :: lb3d_gpu_sanity
  public :: lb3d_gpu_upload_sites, lb3d_gpu_download_sites, lb3d_gpu_init_noise, lb3d_gpu_get_arrows
  public :: lb3d_gpu_bc_init, lb3d_gpu_force_apply, lb3d_gpu_bc_apply
  public :: lb3d_gpu_advection, lb3d_gpu_collision, lb3d_gpu_run_steps
  public :: lb3d_gpu_get_nt, lb3d_gpu_diagnostics, lb3d_gpu_le_state
  public :: lb3d_gpu_synchronize, lb3d_gpu_autotune
  public :: lb3d_gpu_error_message

  interface
     function lb3d_gpu_abi_version() bind(C, name="lb3d_gpu_abi_version") result(v)
       import :: c_int32_t
       integer(c_int32_t) :: v
     end function lb3d_gpu_abi_version

     function c_last_error() bind(C, name="lb3d_gpu_last_error") result(p)
       import :: c_ptr
       type(c_ptr) :: p
     end function c_last_error

     function lb3d_gpu_create(p, ctx) bind(C, name="lb3d_gpu_create") result(rc)
       import :: c_int32_t, c_ptr, lb3d_params
       type(lb3d_params), intent(in) :: p
       type(c_ptr), intent(out) :: ctx
       integer(c_int32_t) :: rc
     end function lb3d_gpu_create

     function lb3d_gpu_destroy(ctx) bind(C, name="lb3d_gpu_destroy") result(rc)
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t) :: rc
     end function lb3d_gpu_destroy

     function lb3d_gpu_upload_sites(ctx, whole_n, halo, sd, nt0) &
          bind(C, name="lb3d_gpu_upload_sites") result(rc)
       import :: c_int32_t, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: whole_n(*)
       integer(c_int32_t), value :: halo, sd, nt0
       integer(c_int32_t) :: rc
     end function lb3d_gpu_upload_sites

     function lb3d_gpu_download_sites(ctx, whole_n, halo, sd) &
          bind(C, name="lb3d_gpu_download_sites") result(rc)
       import :: c_int32_t, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(inout) :: whole_n(*)
       integer(c_int32_t), value :: halo, sd
       integer(c_int32_t) :: rc
     end function lb3d_gpu_download_sites

     function lb3d_gpu_get_arrows(ctx, out) bind(C, name="lb3d_gpu_get_arrows") result(rc)
       import :: c_int32_t, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(inout) :: out(*)
       integer(c_int32_t) :: rc
     end function lb3d_gpu_get_arrows

     function lb3d_gpu_init_noise(ctx, seed, rho_mean, amp, dipoles) &
          bind(C, name="lb3d_gpu_init_noise") result(rc)
       import :: c_int32_t, c_int64_t, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int64_t), value :: seed
       real(c_double), intent(in) :: rho_mean(3), amp(3)
       integer(c_int32_t), value :: dipoles
       integer(c_int32_t) :: rc
     end function lb3d_gpu_init_noise

     function lb3d_gpu_bc_init(ctx) bind(C, name="lb3d_gpu_bc_init") result(rc)
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t) :: rc
     end function lb3d_gpu_bc_init

     function lb3d_gpu_force_apply(ctx, stage) bind(C, name="lb3d_gpu_force_apply") result(rc)
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: stage
       integer(c_int32_t) :: rc
     end function lb3d_gpu_force_apply

     function lb3d_gpu_bc_apply(ctx, stage, nt) bind(C, name="lb3d_gpu_bc_apply") result(rc)
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: stage, nt
       integer(c_int32_t) :: rc
     end function lb3d_gpu_bc_apply

     function lb3d_gpu_advection(ctx) bind(C, name="lb3d_gpu_advection") result(rc)
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t) :: rc
     end function lb3d_gpu_advection

     function lb3d_gpu_collision(ctx) bind(C, name="lb3d_gpu_collision") result(rc)
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t) :: rc
     end function lb3d_gpu_collision

     function lb3d_gpu_run_steps(ctx, nsteps) bind(C, name="lb3d_gpu_run_steps") result(rc)
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: nsteps
       integer(c_int32_t) :: rc
     end function lb3d_gpu_run_steps

     function lb3d_gpu_get_nt(ctx, nt) bind(C, name="lb3d_gpu_get_nt") result(rc)
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(out) :: nt
       integer(c_int32_t) :: rc
     end function lb3d_gpu_get_nt

     function lb3d_gpu_diagnostics(ctx, mass, momentum, counts) &
          bind(C, name="lb3d_gpu_diagnostics") result(rc)
       import :: c_int32_t, c_int64_t, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: mass(3), momentum(3)
       integer(c_int64_t), intent(out) :: counts(3)
       integer(c_int32_t) :: rc
     end function lb3d_gpu_diagnostics

     function lb3d_gpu_sanity(ctx, rep) bind(C, name="lb3d_gpu_sanity") result(rc)
       import :: c_int32_t, c_ptr, lb3d_sanity
       type(c_ptr), value :: ctx
       type(lb3d_sanity), intent(out) :: rep
       integer(c_int32_t) :: rc
     end function lb3d_gpu_sanity

     function lb3d_gpu_le_state(ctx, shear_sum, lelo, lefrac, shear_u) &
          bind(C, name="lb3d_gpu_le_state") result(rc)
       import :: c_int32_t, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: shear_sum, lefrac, shear_u
       integer(c_int32_t), intent(out) :: lelo
       integer(c_int32_t) :: rc
     end function lb3d_gpu_le_state

     function lb3d_gpu_synchronize(ctx) bind(C, name="lb3d_gpu_synchronize") result(rc)
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t) :: rc
     end function lb3d_gpu_synchronize

     function lb3d_gpu_autotune(ctx, trial_steps, report, report_len) &
          bind(C, name="lb3d_gpu_autotune") result(rc)
       import :: c_int32_t, c_ptr, c_char
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: trial_steps
       character(kind=c_char), intent(out) :: report(*)
       integer(c_int32_t), value :: report_len
       integer(c_int32_t) :: rc
     end function lb3d_gpu_autotune

     function c_strlen(s) bind(C, name="strlen") result(n)
       import :: c_ptr, c_size_t
       type(c_ptr), value :: s
       integer(c_size_t) :: n
     end function c_strlen
  end interface

contains

  !> Text of the library's last error on this thread.
  function lb3d_gpu_error_message() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n, i
    p = c_last_error()
    n = int(c_strlen(p))
    allocate(character(len=n) :: msg)
    if (n == 0) return
    call c_f_pointer(p, chars, [n])
    do i = 1, n
       msg(i:i) = chars(i)
    end do
  end function lb3d_gpu_error_message

end module lb3d_gpu_module
