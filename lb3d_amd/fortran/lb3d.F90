!> program lb3d (code/lb3d.F90:29-43): init -> time loop -> finalise, with the
!> lattice update on the GPU.
!>
!> Usage: lb3d -f <input file> [-c 2|3] [-p fp64|fp32] [-n n_iteration]
program lb3d
  use iso_c_binding
  use lb3d_gpu_module
  use lb3d_host_config
  use lb3d_host_state
  use lb3d_host_io
  use lb3d_bc_module
  use lb3d_time_loop_module
  implicit none
  character(len=512) :: inp, arg
  integer :: i, nargs, ov_ncomp, ov_niter
  character(len=16) :: ov_prec

  inp = 'input-file'
  ov_ncomp = -1
  ov_niter = -1
  ov_prec = ''
  nargs = command_argument_count()
  i = 1
  do while (i <= nargs)
     call get_command_argument(i, arg)
     select case (trim(arg))
     case ('-f')
        i = i + 1
        call get_command_argument(i, inp)
     case ('-c')
        i = i + 1
        call get_command_argument(i, arg)
        read(arg, *) ov_ncomp
     case ('-p')
        i = i + 1
        call get_command_argument(i, ov_prec)
     case ('-n')
        i = i + 1
        call get_command_argument(i, arg)
        read(arg, *) ov_niter
     case default
        write(*, '(A,A)') 'unknown argument ', trim(arg)
        error stop 1
     end select
     i = i + 1
  end do

  call lb3d_init()
  call lb3d_time_loop()
  call lb3d_finalise()

contains

  !> Staged initialisation (code/lb3d_initialise.F90:44-97): configuration,
  !> lattice allocation, initial condition, boundary init.
  subroutine lb3d_init()
    type(lb3d_params) :: p
    integer :: c, s
    integer, parameter :: g(19) = [2, 2, 2, 2, 2, 2, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1]
    real(rk) :: rsum
    logical :: walls
    character(len=1024) :: rep
    integer :: k
    call log_msg('lb3d on MI355X: reading '//trim(inp))
    call lb3d_config_read(inp)
    if (ov_ncomp > 0) ncomp = ov_ncomp
    if (ov_niter >= 0) n_iteration = ov_niter
    if (len_trim(ov_prec) > 0) precision = ov_prec
    if (lb3d_gpu_abi_version() /= LB3D_ABI_VERSION) then
       write(*, '(A)') 'FATAL ERROR: liblb3d_gpu ABI version mismatch'
       error stop 1
    end if
    sd = merge(79, 49, ncomp == 3)
    allocate(whole_N(sd, 0:nx + 1, 0:ny + 1, 0:nz + 1))
    whole_N = 0.0_rk
    tstart = 0
    if (init_cond == 7) then
       call read_checkpoint_xdr()
       read(restore_string(2:9), *) tstart
    end if
    nt = tstart
    ! lb3d_ic_init stage 4: the initial condition, then the rock (code/lb3d_ic.F90:76-97)
    walls = lb3d_init_rock()

    p%nx = nx; p%ny = ny; p%nz = nz
    p%ncomp = ncomp
    p%bdist = bdist
    p%precision = merge(LB3D_FP32_DEV, LB3D_FP64, trim(precision) == 'fp32')
    p%boundary = merge(LB3D_BC_LEES_EDWARDS, LB3D_BC_PERIODIC, &
         boundary_cond /= 0 .and. (inv_fluid == 5 .or. inv_fluid == 6))
    p%restore = merge(1, 0, init_cond == 7)
    p%amass = [amass_r, amass_b, amass_s]
    p%tau = [tau_r, tau_b, tau_s]
    p%tau_d = tau_d; p%beta = beta; p%d_0 = d_0
    p%g_br = g_br; p%g_bs = g_bs; p%g_ss = g_ss
    p%shear_u = shear_u; p%shear_omega = shear_omega; p%shear_sum = shear_sum_in
    p%g_accn = [g_accn_x, g_accn_y, g_accn]
    p%g_accn_min = [g_accn_min_x, g_accn_min_y, g_accn_min]
    p%g_accn_max = [g_accn_max_x, g_accn_max_y, g_accn_max]
    p%g_wr = g_wr; p%g_wb = g_wb; p%tau_wr = tau_wr; p%tau_wb = tau_wb
    p%rock_colour_r = rock_colour_r; p%rock_colour_b = rock_colour_b
    ! invasion after the advection (lbe_invade, code/lb3d_bc.F90:129-133): the
    ! Zou-He flux boundary (inv_fluid 12) with (uz_r, uz_g, uz_b) = (pr, pg, pb)
    if (boundary_cond /= 0 .and. inv_fluid >= 0 .and. all(inv_fluid /= [5, 6, 12])) then
       write(*, '(A,I0)') 'FATAL ERROR: this invasion type is not supported: inv_fluid = ', inv_fluid
       error stop 1
    end if
    p%inv_fluid = merge(12, 0, boundary_cond /= 0 .and. inv_fluid == 12)
    p%inv_type = inv_type
    p%inv_flux = [pr, pb, pg]
    ! reference densities of the FP32 deviation form: the mean of the initial state
    p%rho0 = 0.0_c_double
    if (restored) then
       do c = 1, ncomp
          rsum = 0.0_rk
          do s = 1, 19
             rsum = rsum + g(s) * sum(whole_N(24 * (c - 1) + s, 1:nx, 1:ny, 1:nz))
          end do
          p%rho0(c) = rsum / (real(nx, rk) * ny * nz)
       end do
    else
       p%rho0 = [fr, fb, fg]
    end if
    where (rho0 > 0.0d0) p%rho0 = rho0
    where (p%rho0 <= 0.0_c_double) p%rho0 = 0.7_c_double

    call check(lb3d_gpu_create(p, ctx), 'lb3d_gpu_create')
    if (restored) then
       call check(lb3d_gpu_upload_sites(ctx, whole_N, 1_c_int32_t, int(sd, c_int32_t), &
            int(tstart, c_int32_t)), 'upload_sites')
    else
       call check(lb3d_gpu_init_noise(ctx, int(seed, c_int64_t), [fr, fb, fg], [pr, pb, pg], 1_c_int32_t), &
            'init_noise')
       if (walls) then
          ! the rock goes onto the generated state
          call check(lb3d_gpu_download_sites(ctx, whole_N, 1_c_int32_t, int(sd, c_int32_t)), 'download_sites')
          walls = lb3d_init_rock()
          call check(lb3d_gpu_upload_sites(ctx, whole_N, 1_c_int32_t, int(sd, c_int32_t), &
               int(tstart, c_int32_t)), 'upload_sites')
       end if
    end if
    call lb3d_bc_init(4)
    if (autotune) then
       call check(lb3d_gpu_autotune(ctx, 3_c_int32_t, rep, int(len(rep), c_int32_t)), 'autotune')
       k = index(rep, 'chosen:')
       if (k > 0) call log_msg('autotune '//rep(k:min(len(rep), k + 60)))
    end if
    call execute_command_line('mkdir -p '//trim(folder)//'/'//trim(cpfolder))
    call log_msg('init done')
  end subroutine lb3d_init

  !> lb3d_finalise with calculate_performance (code/lb3d_timer.F90:172-195).
  subroutine lb3d_finalise()
    character(len=256) :: msg
    real(rk) :: ups
    if (loop_seconds > 0.0_rk) then
       ups = real(timesteps_count, rk) * real(nx, rk) * real(ny, rk) * real(nz, rk) / loop_seconds
       write(msg, '("Performance: ",ES12.5," site updates per second (",F10.1," MLUPS) over ",I0," steps")') &
            ups, ups / 1.0e6_rk, timesteps_count
       call log_msg(msg)
    end if
    call check(lb3d_gpu_destroy(ctx), 'lb3d_gpu_destroy')
    call log_msg('finished')
  end subroutine lb3d_finalise

end program lb3d
