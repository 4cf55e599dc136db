!> Fortran host of the MI355X LB3D timestep: the reference's time loop and
!> call surface (lb3d_bc_apply, lb3d_advection, lb3d_collision,
!> lb3d_force_apply, lb3d_sanity_check) with the lattice held on the GPU
!> behind the C ABI of include/lb3d_gpu.h.
!>
!> Input is the reference's namelist file (&fixed_input, &variable_input,
!> &lbe_input; code/lb3d_config.F90:45-62, 196-208, 319-335) plus an optional
!> &gpu_input group (ncomp, precision, autotune, fast_loop).  State in and
!> out uses the reference's XDR checkpoint format and observable dumps
!> (code/lb3d_io_checkpoint.F90:741-1106, code/lb3d_io_xdrf.F90:717-912).

!------------------------------------------------------------------------------
module lb3d_host_config
  implicit none
  integer, parameter :: rk = 8

  ! &fixed_input
  integer :: nx = 16, ny = 16, nz = 16, seed = 1, boundary_cond = 0
  character(len=256) :: obs_file = 'empty.dat', obs_folder = '.'
  namelist /fixed_input/ nx, ny, nz, seed, obs_file, obs_folder, boundary_cond

  ! &variable_input
  integer :: n_iteration = 100, n_sci_start = 0, init_cond = 0, inv_fluid = -1, inv_type = -1
  logical :: sci_int = .false., sci_sur = .false., sci_od = .false., sci_wd = .false.
  logical :: sci_dir = .false., sci_vel = .false., sci_flo = .false., sci_arrows = .false.
  logical :: sci_pressure = .false., sci_velocities = .false., sci_rock = .false.
  logical :: sci_fluxz = .false., sci_profile = .false., post = .false.
  integer :: n_sci_int = 1, n_sci_sur = 1, n_sci_od = 1, n_sci_wd = 1, n_sci_dir = 1
  integer :: n_sci_vel = 1, n_sci_flo = 1, n_sci_arrows = 1, n_sci_pressure = 1
  integer :: n_sci_velocities = 1, n_sci_rock = 1, n_sci_fluxz = 1, n_sci_profile = 1
  integer :: sci_pressure_init = 0
  character(len=256) :: folder = 'output', cpfolder = '.', srccpfolder = '', gr_out_file = 'lb3d'
  real(rk) :: fr = 0.7_rk, fg = 0.7_rk, fb = 0.7_rk, fd = 0.0_rk, fr1 = 0.0_rk, fr2 = 0.0_rk
  real(rk) :: pr = 0.01_rk, pg = 0.0_rk, pb = 0.01_rk, rock_colour = 0.0_rk, beta = 1.0_rk
  real(rk) :: rock_colour_r = 0.0_rk, rock_colour_b = 0.0_rk
  real(rk) :: drop_xshift = 0.0_rk, drop_yshift = 0.0_rk, drop_zshift = 0.0_rk
  namelist /variable_input/ n_iteration, n_sci_start, sci_int, n_sci_int, sci_sur, n_sci_sur, &
       sci_od, n_sci_od, sci_wd, n_sci_wd, sci_dir, n_sci_dir, sci_vel, n_sci_vel, sci_flo, &
       n_sci_flo, sci_arrows, n_sci_arrows, sci_pressure, n_sci_pressure, sci_pressure_init, &
       sci_velocities, n_sci_velocities, sci_rock, n_sci_rock, sci_fluxz, n_sci_fluxz, &
       sci_profile, n_sci_profile, post, folder, cpfolder, srccpfolder, gr_out_file, init_cond, &
       fr, fg, fb, fd, fr1, fr2, pr, pb, pg, inv_fluid, inv_type, rock_colour, rock_colour_r, &
       rock_colour_b, beta, drop_xshift, drop_yshift, drop_zshift

  ! &lbe_input (defaults of code/lb3d_config.F90:244-297)
  real(rk) :: amass_r = 1.0_rk, amass_b = 1.0_rk, amass_s = 1.0_rk
  real(rk) :: tau_r = 1.0_rk, tau_b = 1.0_rk, tau_s = 1.0_rk, tau_d = 2.0_rk
  real(rk) :: g_br = 0.0_rk, g_bs = 0.0_rk, g_ss = 0.0_rk, g_wr = 0.0_rk, d_0 = 1.0_rk
  real(rk) :: g_wb = 0.0_rk, tau_wr = 1.0_rk, tau_wb = 1.0_rk
  real(rk) :: perturbation = 0.0_rk, shear_u = 0.0_rk, shear_omega = 0.0_rk, acccoef = 0.0_rk
  real(rk) :: g_accn = 0.0_rk, g_accn_x = 0.0_rk, g_accn_y = 0.0_rk
  integer :: g_accn_min = 0, g_accn_max = 0, g_accn_min_x = 0, g_accn_max_x = 0
  integer :: g_accn_min_y = 0, g_accn_max_y = 0
  integer :: bdist = 0, psifunc = 2, bcsel = 0, n_checkpoint = 300, num_chkp_files = 2
  integer :: n_sanity_check = 1000
  logical :: checkpoint_safe = .true., write_AVS_fld = .false., dump_double = .false.
  logical :: SCMP = .false., OXFORD = .false.
  character(len=64) :: restore_string = 't00000000-0000000000', dump_format = 'xdr'
  character(len=64) :: checkpoint_format = 'xdr'
  namelist /lbe_input/ amass_b, amass_r, amass_s, tau_b, tau_r, tau_s, tau_d, g_br, g_bs, g_ss, &
       g_wr, g_wb, tau_wr, tau_wb, d_0, bdist, perturbation, shear_u, shear_omega, g_accn, &
       g_accn_x, g_accn_y, &
       g_accn_min, g_accn_max, g_accn_min_x, g_accn_max_x, g_accn_min_y, g_accn_max_y, &
       n_checkpoint, restore_string, num_chkp_files, psifunc, dump_format, checkpoint_safe, &
       checkpoint_format, write_AVS_fld, dump_double, n_sanity_check, SCMP, OXFORD, bcsel, acccoef

  ! &gpu_input (this host's own group; optional)
  integer :: ncomp = 3                  ! 3 = the reference's default build, 2 = -DNOSURFACTANT
  character(len=16) :: precision = 'fp64'
  logical :: autotune = .false.
  logical :: fast_loop = .true.         ! replay captured steps between I/O points
  ! reference densities of the FP32 deviation form; <= 0: the mean of the
  ! initial (or restored) state.  Fixing them makes FP32 restarts bitwise.
  real(8) :: rho0(3) = (/0.0d0, 0.0d0, 0.0d0/)
  namelist /gpu_input/ ncomp, precision, autotune, fast_loop, rho0

contains

  !> Read the namelist file (the reference reads the same groups, code/lb3d_config.F90:340-371).
  subroutine lb3d_config_read(fname)
    character(len=*), intent(in) :: fname
    integer :: u, ios
    open(newunit=u, file=fname, status='old', action='read', iostat=ios)
    if (ios /= 0) then
       write(*, '(A,A)') 'FATAL ERROR: cannot open input file ', trim(fname)
       error stop 1
    end if
    read(u, nml=fixed_input, iostat=ios)
    if (ios /= 0) call nml_fail('fixed_input', ios)
    rewind(u)
    read(u, nml=variable_input, iostat=ios)
    if (ios /= 0) call nml_fail('variable_input', ios)
    rewind(u)
    read(u, nml=lbe_input, iostat=ios)
    if (ios /= 0) call nml_fail('lbe_input', ios)
    rewind(u)
    read(u, nml=gpu_input, iostat=ios)   ! optional group
    close(u)
  end subroutine lb3d_config_read

  subroutine nml_fail(group, ios)
    character(len=*), intent(in) :: group
    integer, intent(in) :: ios
    write(*, '(A,A,A,I0)') 'FATAL ERROR: reading namelist &', group, ', iostat ', ios
    error stop 1
  end subroutine nml_fail

end module lb3d_host_config

!------------------------------------------------------------------------------
module lb3d_host_state
  use iso_c_binding
  use lb3d_gpu_module
  use lb3d_host_config, only: rk
  implicit none
  type(c_ptr) :: ctx = c_null_ptr
  real(rk), allocatable, target :: whole_N(:, :, :, :)   ! (site_doubles, 0:nx+1, 0:ny+1, 0:nz+1)
  integer :: sd = 79
  integer :: nt = 0, tstart = 0, timesteps_count = 0
  real(rk) :: shear_sum_in = 0.0_rk
  logical :: restored = .false.

contains

  !> Abort on a library error (the reference calls abend -> MPI_Abort,
  !> code/lb3d_mpi.F90:670-676).
  subroutine check(rc, where)
    integer(c_int32_t), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc /= LB3D_OK) then
       write(*, '(A,A,A,I0,A,A)') 'FATAL ERROR in ', where, ': code ', rc, ': ', lb3d_gpu_error_message()
       error stop 1
    end if
  end subroutine check

  subroutine log_msg(msg)
    character(len=*), intent(in) :: msg
    integer :: v(8)
    call date_and_time(values=v)
    write(*, '(I2.2,":",I2.2,":",I2.2," - ",A)') v(5), v(6), v(7), trim(msg)
  end subroutine log_msg

end module lb3d_host_state

!------------------------------------------------------------------------------
!> Big-endian XDR files of the reference (checkpoints and float32 dumps).
module lb3d_host_io
  use iso_c_binding
  use lb3d_host_config
  use lb3d_host_state
  implicit none

contains

  subroutine swap8(b)
    integer(1), intent(inout) :: b(:)
    integer :: i, k
    integer(1) :: t
    do i = 1, size(b), 8
       do k = 0, 3
          t = b(i + k)
          b(i + k) = b(i + 7 - k)
          b(i + 7 - k) = t
       end do
    end do
  end subroutine swap8

  subroutine swap4(b)
    integer(1), intent(inout) :: b(:)
    integer :: i
    integer(1) :: t
    do i = 1, size(b), 4
       t = b(i); b(i) = b(i + 3); b(i + 3) = t
       t = b(i + 1); b(i + 1) = b(i + 2); b(i + 2) = t
    end do
  end subroutine swap4

  !> Per-site record of the checkpoint: n_r(19) n_b(19) [n_s(19) da(3) db(3)] rock
  !> (code/lb3d_io_checkpoint.F90:833-871); returns the lbe_site offsets (1-based).
  function record_offsets() result(off)
    integer, allocatable :: off(:)
    integer :: i, k
    if (ncomp == 3) then
       allocate(off(64))
       k = 0
       do i = 1, 19; k = k + 1; off(k) = i; end do
       do i = 1, 19; k = k + 1; off(k) = 24 + i; end do
       do i = 1, 19; k = k + 1; off(k) = 48 + i; end do
       do i = 1, 3; k = k + 1; off(k) = 72 + i; end do
       do i = 1, 3; k = k + 1; off(k) = 75 + i; end do
       off(64) = 79
    else
       allocate(off(39))
       k = 0
       do i = 1, 19; k = k + 1; off(k) = i; end do
       do i = 1, 19; k = k + 1; off(k) = 24 + i; end do
       off(39) = 49
    end if
  end function record_offsets

  !> Restore whole_N from the reference's XDR checkpoint written by one rank.
  subroutine read_checkpoint_xdr()
    character(len=512) :: fname
    integer :: u, ios, x, y, z, rec, k
    integer, allocatable :: off(:)
    integer(1), allocatable :: bytes(:)
    real(rk), allocatable :: vals(:)
    integer(1) :: topo(32)
    off = record_offsets()
    rec = size(off)
    write(fname, '(A,"/checkpoint_",A,"_",A,"_p000000.xdr")') trim(srccpfolder), trim(gr_out_file), &
         trim(restore_string)
    call log_msg('Restoring checkpoint '//trim(fname))
    open(newunit=u, file=fname, access='stream', form='unformatted', status='old', action='read', iostat=ios)
    if (ios /= 0) then
       write(*, '(A,A)') 'FATAL ERROR: cannot open checkpoint ', trim(fname)
       error stop 1
    end if
    allocate(bytes(8 * rec * nx * ny * nz), vals(rec * nx * ny * nz))
    read(u, iostat=ios) bytes
    close(u)
    if (ios /= 0) then
       write(*, '(A)') 'FATAL ERROR: short checkpoint file'
       error stop 1
    end if
    call swap8(bytes)
    vals = transfer(bytes, vals)
    whole_N = 0.0_rk
    k = 0
    do x = 1, nx
       do y = 1, ny
          do z = 1, nz
             whole_N(off, x, y, z) = vals(k + 1:k + rec)
             k = k + rec
          end do
       end do
    end do
    ! checktopo: cdims(3), ccoords(3) int32 and shear_sum f64 (code/lb3d_io_checkpoint.F90:763-774)
    write(fname, '(A,"/checktopo_",A,"_",A,".xdr")') trim(srccpfolder), trim(gr_out_file), trim(restore_string)
    open(newunit=u, file=fname, access='stream', form='unformatted', status='old', action='read', iostat=ios)
    shear_sum_in = 0.0_rk
    if (ios == 0) then
       read(u, iostat=ios) topo
       close(u)
       if (ios == 0) then
          call swap8(topo(25:32))
          shear_sum_in = transfer(topo(25:32), shear_sum_in)
       end if
    end if
    restored = .true.
  end subroutine read_checkpoint_xdr

  !> lbe_init_rock (code/lb3d_ic_rock.F90:50-189, one rank, no XDRROCKWET):
  !> read obs_folder/obs_file (one big-endian float32 per site, z outer ..
  !> x inner, read_rock_xdrf_par, code/lb3d_io_xdrf.F90:400-470): a site is
  !> rock where the value exceeds 0.  Rock sites lose their populations (and
  !> ternary da, db) and take the positive rock colours.  Returns whether any
  !> rock was set.
  logical function lb3d_init_rock() result(any_rock)
    character(len=512) :: fname
    integer :: u, ios, x, y, z, c
    integer(1), allocatable :: bytes(:)
    real(4), allocatable :: vals(:)
    integer :: k
    any_rock = .false.
    ! rock_colour is the deprecated form of rock_colour_r / rock_colour_b
    ! (code/lb3d_config.F90:657-665)
    if (rock_colour > 0) then
       rock_colour_r = rock_colour
    else if (rock_colour < 0) then
       rock_colour_b = rock_colour
    end if
    if (trim(obs_file) == 'empty.dat') return
    fname = trim(obs_folder)//'/'//trim(obs_file)
    if (index(fname, '.xdr') == 0) then
       write(*, '(A,A)') 'FATAL ERROR: only XDR obstacle files are read: ', trim(fname)
       error stop 1
    end if
    call log_msg('Reading rock data file <'//trim(fname)//'>')
    open(newunit=u, file=fname, access='stream', form='unformatted', status='old', action='read', iostat=ios)
    if (ios /= 0) then
       write(*, '(A,A)') 'FATAL ERROR: cannot open obstacle file ', trim(fname)
       error stop 1
    end if
    allocate(bytes(4 * nx * ny * nz), vals(nx * ny * nz))
    read(u, iostat=ios) bytes
    close(u)
    if (ios /= 0) then
       write(*, '(A)') 'FATAL ERROR: short obstacle file'
       error stop 1
    end if
    call swap4(bytes)
    vals = transfer(bytes, vals)
    k = 0
    do z = 1, nz
       do y = 1, ny
          do x = 1, nx
             k = k + 1
             if (vals(k) > 0.0) then
                whole_N(sd, x, y, z) = 1.0_rk
                do c = 1, ncomp
                   whole_N(24 * (c - 1) + 1:24 * (c - 1) + 19, x, y, z) = 0.0_rk
                end do
                if (ncomp == 3) whole_N(73:78, x, y, z) = 0.0_rk
                if (rock_colour_r > 0) whole_N(24, x, y, z) = rock_colour_r
                if (rock_colour_b > 0) whole_N(48, x, y, z) = rock_colour_b
                any_rock = .true.
             else
                whole_N(sd, x, y, z) = 0.0_rk
             end if
          end do
       end do
    end do
  end function lb3d_init_rock

  !> Write whole_N in the reference's XDR checkpoint layout (one rank).
  subroutine write_checkpoint_xdr()
    character(len=512) :: fname
    integer :: u, x, y, z, rec, k
    integer(1) :: topo(32)
    real(c_double) :: ss, lefrac, su
    integer(c_int32_t) :: lelo
    integer, allocatable :: off(:)
    integer(1), allocatable :: bytes(:)
    real(rk), allocatable :: vals(:)
    off = record_offsets()
    rec = size(off)
    allocate(vals(rec * nx * ny * nz))
    k = 0
    do x = 1, nx
       do y = 1, ny
          do z = 1, nz
             vals(k + 1:k + rec) = whole_N(off, x, y, z)
             k = k + rec
          end do
       end do
    end do
    bytes = transfer(vals, bytes)
    call swap8(bytes)
    write(fname, '(A,"/",A,"/checkpoint_",A,"_t",I8.8,"-",I10.10,"_p000000.xdr")') trim(folder), &
         trim(cpfolder), trim(gr_out_file), nt, 0
    open(newunit=u, file=fname, access='stream', form='unformatted', status='replace', action='write')
    write(u) bytes
    close(u)
    call log_msg('Wrote checkpoint '//trim(fname))
    ! checktopo: cdims(3) = 1, ccoords(3) = 0 (one rank) and shear_sum
    ! (dump_lattice_MPrxdr, code/lb3d_io_checkpoint.F90:763-774)
    call check(lb3d_gpu_le_state(ctx, ss, lelo, lefrac, su), 'le_state')
    topo(1:24) = transfer((/1_4, 1_4, 1_4, 0_4, 0_4, 0_4/), topo(1:24))
    call swap4(topo(1:24))
    topo(25:32) = transfer(ss, topo(25:32))
    call swap8(topo(25:32))
    write(fname, '(A,"/",A,"/checktopo_",A,"_t",I8.8,"-",I10.10,".xdr")') trim(folder), &
         trim(cpfolder), trim(gr_out_file), nt, 0
    open(newunit=u, file=fname, access='stream', form='unformatted', status='replace', action='write')
    write(u) topo
    close(u)
  end subroutine write_checkpoint_xdr

  subroutine dump_scalar_xdr(name, field)
    character(len=*), intent(in) :: name
    real(rk), intent(in) :: field(:, :, :)
    character(len=512) :: fname
    real(4), allocatable :: f4(:)
    integer(1), allocatable :: bytes(:)
    integer :: u
    f4 = reshape(real(field, 4), [size(field)])
    bytes = transfer(f4, bytes)
    call swap4(bytes)
    write(fname, '(A,"/",A,"_",A,"_t",I8.8,"-",I10.10,".xdr")') trim(folder), name, trim(gr_out_file), nt, 0
    open(newunit=u, file=fname, access='stream', form='unformatted', status='replace', action='write')
    write(u) bytes
    close(u)
  end subroutine dump_scalar_xdr

  subroutine dump_vector_xdr(name, vec)
    character(len=*), intent(in) :: name
    real(rk), intent(in) :: vec(:, :, :, :)   ! (3, nx, ny, nz)
    call dump_scalar_xdr_raw(name, reshape(real(vec, 4), [size(vec)]))
  end subroutine dump_vector_xdr

  subroutine dump_scalar_xdr_raw(name, f4)
    character(len=*), intent(in) :: name
    real(4), intent(in) :: f4(:)
    character(len=512) :: fname
    integer(1), allocatable :: bytes(:)
    integer :: u
    bytes = transfer(f4, bytes)
    call swap4(bytes)
    write(fname, '(A,"/",A,"_",A,"_t",I8.8,"-",I10.10,".xdr")') trim(folder), name, trim(gr_out_file), nt, 0
    open(newunit=u, file=fname, access='stream', form='unformatted', status='replace', action='write')
    write(u) bytes
    close(u)
  end subroutine dump_scalar_xdr_raw

  logical function dump_now(flag, every)
    logical, intent(in) :: flag
    integer, intent(in) :: every
    dump_now = flag .and. every > 0 .and. nt >= n_sci_start
    if (dump_now) dump_now = mod(nt, every) == 0
  end function dump_now

  !> The observables od, wd, sur, dir, vel of the reference's dumps
  !> (code/lb3d_io_dump_data.F90:922-1166) from the downloaded lattice.
  subroutine lb3d_io_write_data()
    integer, parameter :: g(19) = [2, 2, 2, 2, 2, 2, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1]
    integer, parameter :: cz(19) = [0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, 0]
    real(rk), allocatable :: sc(:, :, :), vec(:, :, :, :)
    real(rk) :: amass(3), omega(3), rho, mom, den, v
    integer :: x, y, z, c, s
    logical :: any_dump
    any_dump = dump_now(sci_od, n_sci_od) .or. dump_now(sci_wd, n_sci_wd) .or. &
         dump_now(sci_sur, n_sci_sur) .or. dump_now(sci_dir, n_sci_dir) .or. dump_now(sci_vel, n_sci_vel) .or. &
         dump_now(sci_arrows, n_sci_arrows)
    if (.not. any_dump) return
    if (dump_now(sci_arrows, n_sci_arrows)) call dump_arrows()
    call check(lb3d_gpu_download_sites(ctx, whole_N, 1_c_int32_t, int(sd, c_int32_t)), 'download_sites')
    amass = [amass_r, amass_b, amass_s]
    omega = 1.0_rk / [tau_r, tau_b, tau_s]
    allocate(sc(nx, ny, nz))
    do c = 1, min(ncomp, 3)
       if (c == 1 .and. .not. dump_now(sci_od, n_sci_od)) cycle
       if (c == 2 .and. .not. dump_now(sci_wd, n_sci_wd)) cycle
       if (c == 3 .and. .not. dump_now(sci_sur, n_sci_sur)) cycle
       do z = 1, nz
          do y = 1, ny
             do x = 1, nx
                rho = 0.0_rk
                do s = 1, 19
                   rho = rho + whole_N(24 * (c - 1) + s, x, y, z) * g(s)
                end do
                sc(x, y, z) = amass(c) * rho
                ! od and wd are 0 on rock sites (dump_od/dump_wd, code/lb3d_io_dump_data.F90:1120, 1156)
                if (c < 3 .and. whole_N(sd, x, y, z) /= 0.0_rk) sc(x, y, z) = 0.0_rk
             end do
          end do
       end do
       if (c == 1) call dump_scalar_xdr('od', sc)
       if (c == 2) call dump_scalar_xdr('wd', sc)
       if (c == 3) call dump_scalar_xdr('sur', sc)
    end do
    if (ncomp == 3 .and. dump_now(sci_dir, n_sci_dir)) then
       allocate(vec(3, nx, ny, nz))
       vec = whole_N(76:78, 1:nx, 1:ny, 1:nz)
       call dump_vector_xdr('dir', vec)
    end if
    if (dump_now(sci_vel, n_sci_vel)) then
       do z = 1, nz
          do y = 1, ny
             do x = 1, nx
                v = 0.0_rk
                den = 0.0_rk
                do c = 1, ncomp
                   mom = 0.0_rk
                   rho = 0.0_rk
                   do s = 1, 19
                      mom = mom + whole_N(24 * (c - 1) + s, x, y, z) * g(s) * cz(s)
                      rho = rho + whole_N(24 * (c - 1) + s, x, y, z) * g(s)
                   end do
                   v = v + mom * amass(c) * omega(c)
                   den = den + rho * amass(c) * omega(c)
                end do
                v = v / max(10.d-9, den)
                if (whole_N(sd, x, y, z) /= 0.0_rk) then   ! rock: 0 (dump_vel, :1031, 1086)
                   sc(x, y, z) = 0.0_rk
                   cycle
                end if
                if (z >= g_accn_min .and. z <= merge(nz, g_accn_max, g_accn_max == 0) .and. &
                     x >= g_accn_min_x .and. x <= merge(nx, g_accn_max_x, g_accn_max_x == 0) .and. &
                     y >= g_accn_min_y .and. y <= merge(ny, g_accn_max_y, g_accn_max_y == 0)) &
                     v = v - g_accn / 2.0d0
                sc(x, y, z) = v
             end do
          end do
       end do
       call dump_scalar_xdr('vel', sc)
    end if
  end subroutine lb3d_io_write_data

  !> arr (dump_arrows, code/lb3d_io_dump_data.F90:1308-1460): the velocity
  !> with half the Shan-Chen force taken off, from the library
  !> (lb3d_gpu_get_arrows).  Its forces need the densities of an advection,
  !> so before the first timestep of a run there is nothing to dump: that is
  !> logged, not silently skipped.
  subroutine dump_arrows()
    real(rk), allocatable :: vec(:, :, :, :)
    integer(c_int32_t) :: rc
    allocate(vec(3, nx, ny, nz))
    rc = lb3d_gpu_get_arrows(ctx, vec)
    if (rc == LB3D_ESTATE) then
       call log_msg('WARNING: arr not dumped at this timestep: no advection has run yet')
       return
    end if
    call check(rc, 'get_arrows')
    call dump_vector_xdr('arr', vec)
  end subroutine dump_arrows

  subroutine lb3d_io_write_checkpoint()
    if (n_checkpoint <= 0 .or. nt == tstart) return
    if (mod(nt, n_checkpoint) /= 0) return
    call check(lb3d_gpu_download_sites(ctx, whole_N, 1_c_int32_t, int(sd, c_int32_t)), 'download_sites')
    call write_checkpoint_xdr()
  end subroutine lb3d_io_write_checkpoint

end module lb3d_host_io

!------------------------------------------------------------------------------
!> lb3d_force_apply (code/lb3d_force.F90:98-121): no external force on this path.
module lb3d_force_module
  use lb3d_host_state
  implicit none
contains
  subroutine lb3d_force_apply(stage)
    integer, intent(in) :: stage
    call check(lb3d_gpu_force_apply(ctx, int(stage, c_int32_t)), 'lb3d_force_apply')
  end subroutine lb3d_force_apply
end module lb3d_force_module

!------------------------------------------------------------------------------
!> lb3d_bc_init / lb3d_bc_apply (code/lb3d_bc.F90:48-155): periodic wraps are
!> inside the kernels; Lees-Edwards runs the face passes on the device.
module lb3d_bc_module
  use lb3d_host_state
  implicit none
contains
  subroutine lb3d_bc_init(stage)
    integer, intent(in) :: stage
    if (stage == 4) call check(lb3d_gpu_bc_init(ctx), 'lb3d_bc_init')
  end subroutine lb3d_bc_init

  subroutine lb3d_bc_apply(stage)
    integer, intent(in) :: stage
    call check(lb3d_gpu_bc_apply(ctx, int(stage, c_int32_t), int(nt, c_int32_t)), 'lb3d_bc_apply')
  end subroutine lb3d_bc_apply
end module lb3d_bc_module

!------------------------------------------------------------------------------
!> lb3d_advection (code/lb3d_advection.F90:40-50).
module lb3d_advection_module
  use lb3d_host_state
  implicit none
contains
  subroutine lb3d_advection()
    call check(lb3d_gpu_advection(ctx), 'lb3d_advection')
  end subroutine lb3d_advection
end module lb3d_advection_module

!------------------------------------------------------------------------------
!> lb3d_collision (code/lb3d_collision.F90:41-47).
module lb3d_collision_module
  use lb3d_host_state
  implicit none
contains
  subroutine lb3d_collision()
    call check(lb3d_gpu_collision(ctx), 'lb3d_collision')
  end subroutine lb3d_collision
end module lb3d_collision_module

!------------------------------------------------------------------------------
!> lb3d_sanity_check (code/lb3d_sanity.F90:45-61): FP64 device reductions;
!> a NaN aborts after writing the data, like the reference.
module lb3d_sanity_module
  use lb3d_host_config
  use lb3d_host_state
  use lb3d_host_io
  implicit none
contains
  !> lbe_sanity_check (code/lb3d_sanity.F90:79-384): the same report, from
  !> the device reductions of lb3d_gpu_sanity.
  subroutine lb3d_sanity_check()
    type(lb3d_sanity) :: r
    character(len=256) :: msg
    character(len=4), parameter :: spec_name(3) = (/' red', 'blue', 'surf'/)
    character(len=17), parameter :: avg_name(3) = (/'      red (x,y,z)', '     blue (x,y,z)', &
         '     surf (x,y,z)'/)
    real(rk) :: tot, sumdens, vmax(3)
    integer :: i, nspec
    if (n_sanity_check <= 0) return
    if (mod(nt, n_sanity_check) /= 0) return
    call check(lb3d_gpu_sanity(ctx, r), 'lb3d_sanity_check')
    nspec = ncomp
    tot = real(nx, rk) * real(ny, rk) * real(nz, rk)
    call log_msg('---------( Start LBE sanity check )----------')
    if (r%nan_sites > 0) call log_msg('Found NaN')
    write(msg, "('Total densities (r,b,g)   : ',F16.4,' ',F16.4,' ',F16.4)") &
         max(0.d-30, r%mass(1)), max(0.d-15, r%mass(2)), max(0.d-30, r%mass(3))
    call log_msg(msg)
    write(msg, "('Average densities (r,b,g) : ',F16.10,' ',F16.10,' ',F16.10)") &
         max(0.d-30, r%mass(1)/tot), max(0.d-30, r%mass(2)/tot), max(0.d-30, r%mass(3)/tot)
    call log_msg(msg)
    sumdens = r%mass(1) + r%mass(2) + r%mass(3)
    write(msg, "('Total/average density     : ',F16.4,' ',F16.10)") sumdens, max(0.d-30, sumdens/tot)
    call log_msg(msg)
    call log_msg('Maximum absolute velocity : ')
    do i = 1, nspec
       vmax = r%max_vel(:, i)
       write(msg, "('     ',A4,' (x,y,z)         : ',3(F16.10,:,' '))") spec_name(i), max(0.0e-30_rk, vmax)
       call log_msg(msg)
       write(msg, "('      at position',9X,': ',3('(',2(I4,','),I4,')',:,X))") &
            r%max_pos(:, 1, i), r%max_pos(:, 2, i), r%max_pos(:, 3, i)
       call log_msg(msg)
    end do
    call log_msg('Average velocity          : ')
    do i = 1, nspec
       write(msg, "(A17,'         : ',F16.10,' ',F16.10,' ',F16.10)") avg_name(i), r%vel_sum(:, i) / tot
       call log_msg(msg)
    end do
    do i = 1, nspec
       if (any(r%max_vel(:, i) > 0.1_rk)) &
            call log_msg('********  WARNING: MAXIMUM '//trim(spec_name(i))//' VELOCITY IS TOO HIGH!  ********')
    end do
    if (r%fast_sites /= 0) then
       write(msg, "('********  WARNING: absolute species velocity higher than ',F16.12,' on ',F16.2," &
            //"' lattice sites ( ',F16.12,' % of total fluid sites)  ********')") &
            0.1_rk, real(r%fast_sites, rk), 100.0_rk * real(r%fast_sites, rk) / real(r%fluid_sites, rk)
       call log_msg(msg)
    end if
    if (r%negative_sites /= 0) then
       write(msg, "('********  WARNING: negative distribution functions on ',F16.2," &
            //"' lattice sites ( ',F16.12,' % of total fluid sites)  ********')") &
            real(r%negative_sites, rk), 100.0_rk * real(r%negative_sites, rk) / real(r%fluid_sites, rk)
       call log_msg(msg)
    end if
    write(msg, "('Total momentum (x,y,z)    : ',3(F16.8,:,' '))") r%momentum
    call log_msg(msg)
    call log_msg('----------( End LBE sanity check )-----------')
    if (r%nan_sites > 0) then
       ! lb3d_time_loop's insanity branch: dump, checkpoint and stop
       call lb3d_io_write_data()
       call lb3d_io_write_checkpoint()
       error stop 2
    end if
  end subroutine lb3d_sanity_check
end module lb3d_sanity_module

!------------------------------------------------------------------------------
!> lb3d_time_loop (code/lb3d_time_loop.F90:42-97): the same stages per step.
!> With fast_loop the steps between two I/O or sanity points are replayed
!> from the library's captured HIP graph instead of stage by stage.
module lb3d_time_loop_module
  use lb3d_host_config
  use lb3d_host_state
  use lb3d_host_io
  use lb3d_force_module
  use lb3d_bc_module
  use lb3d_advection_module
  use lb3d_collision_module
  use lb3d_sanity_module
  implicit none
  real(rk) :: loop_seconds = 0.0_rk
contains
  integer function next_event(t) result(e)
    integer, intent(in) :: t
    e = n_iteration
    call nearer(sci_od, n_sci_od)
    call nearer(sci_wd, n_sci_wd)
    call nearer(sci_sur, n_sci_sur)
    call nearer(sci_dir, n_sci_dir)
    call nearer(sci_vel, n_sci_vel)
    call nearer(sci_arrows, n_sci_arrows)
    call nearer(.true., n_checkpoint)
    call nearer(.true., n_sanity_check)
  contains
    subroutine nearer(flag, every)
      logical, intent(in) :: flag
      integer, intent(in) :: every
      integer :: k
      if (.not. flag .or. every <= 0) return
      k = (t / every + 1) * every
      if (k < e) e = k
    end subroutine nearer
  end function next_event

  subroutine lb3d_time_loop()
    integer :: stage, target
    integer(8) :: c0, c1, rate
    character(len=64) :: msgstr
    call lb3d_io_write_data()
    call lb3d_sanity_check()
    timesteps_count = 0
    call system_clock(c0, rate)
    do while (nt < n_iteration)
       if (fast_loop) then
          target = next_event(nt)
          call check(lb3d_gpu_run_steps(ctx, int(target - nt, c_int32_t)), 'lb3d_time_loop')
          timesteps_count = timesteps_count + (target - nt)
          nt = target
       else
          nt = nt + 1
          write(msgstr, "('Starting timestep ',i0)") nt
          call log_msg(msgstr)
          timesteps_count = timesteps_count + 1
          stage = 6
          call lb3d_force_apply(stage)
          call lb3d_bc_apply(stage)
          stage = 7
          call lb3d_advection()
          stage = 8
          call lb3d_bc_apply(stage)
          call lb3d_force_apply(stage)
          stage = 9
          call lb3d_collision()
       end if
       stage = 10
       call lb3d_io_write_data()
       call lb3d_io_write_checkpoint()
       call lb3d_sanity_check()
    end do
    call check(lb3d_gpu_synchronize(ctx), 'lb3d_time_loop')
    call system_clock(c1)
    loop_seconds = real(c1 - c0, rk) / real(rate, rk)
  end subroutine lb3d_time_loop
end module lb3d_time_loop_module
