! axion_3d on one MI355X: the reference's driver (main.f90), parameter keys (settings.f90
! 112-297) and checkpoint schedule (evolve.f90 26-152, 154-235, 401-546), with the lattice and
! its diagnostics resident on the GPU behind the C-ABI of libaxgrav (module axgrav_capi).
!
!   axion_3d <param.ini> [-n tag] [-t hours] [-b GB]
!
! Output files keep the reference names and layouts: <rootname>_gravspec_<idx>.dat
! '(5E20.10e3)', _spectrum_<idx>.dat '(4E20.10e3)', _mask_<idx>-0.dat '(3I6)',
! _string_wall_parameters.txt '(I12,11E20.10e3)', _crho.txt '(I12,3E20.10e3)', the resume
! file _param.chk ('idx tau flag') and _phi_<idx>.chk (raw FP64 complex phi then pi).
!
! Keys this build adds (all optional): precision = fp64 | fp32c, num_bin_grav (25),
! gw_fix_swap (F), dtau_fixed (0: use get_dtau), potential_kappa0 (0), ic_seed (seed of the
! flat-spectrum field of a fresh start with initial_loop = T, up to ic_kmax; the thermal state of
! initial_loop = F draws from the reference's stream seeded by ss, like the string fits and the
! covariance), ic_file (raw phi,pi file to start from instead), chk_gw (F: also checkpoint the GW
! accumulators, which the reference does not), device (0).
program axion_3d
  use, intrinsic :: iso_c_binding
  use axgrav_capi
  use axg_ini
  implicit none

  double precision, parameter :: pi_t = 3.14159265359d0, twopi_t = 2*pi_t  ! potential.f90:8
  double precision, parameter :: mstar_in_hz = 2.435d18*1.5192488833521d24  ! potential.f90:14-23
  double precision, parameter :: byte_dc = 16d0, hard_limit_cdata = 2*byte_dc  ! settings.f90:29-31

  character(len=1024) :: paramfile, tag, arg, rootdir, rootname, tmprn, tmpdir, file_resume, ss, prec
  character(len=1024) :: ic_file
  integer :: nargs, i, ios, n_chk, n_id, num_bin, num_bin_grav, degrade_spec, n_mod_eps, idx, info
  integer(c_int64_t) :: ngrid, n3, ic_seed
  logical :: dry_run, initial_loop, id_strings, id_walls, spec_est, grav_spec, group_strings, group_walls
  logical :: out_strings, out_walls, ok, chk, flag_defects, gw_fix_swap, chk_gw, exists
  double precision :: tsim_max, limit_memory, b, sigma, g_rel, kappa, tau_ini, tau_fin, eps_dtau, boxsize
  double precision :: zeta, width_cr, thickness_cr, tau0, dtau_chk, tau_start, tau_end, ic_kmax
  double precision, allocatable :: tau_id(:), crho_spec(:, :)
  real :: tsim_begin, tsim_now
  type(axg_keys) :: keys
  type(axg_params) :: prm
  type(c_ptr) :: h
  integer(c_int32_t) :: nsteps
  integer(c_int32_t) :: tune_code   ! the launch configuration axg_tune picks
  integer(c_int32_t) :: seed(2)   ! the two integers of key ss (subroutines.f90:49-95)
  integer(8) :: clk
  integer(8) :: clock0, clock_rate

  ! ------------------------------------------------------------ command line (main.f90:16-33)
  nargs = command_argument_count()
  if (nargs > 7) call die('error: too many arguments')
  if (nargs < 1) call die('error: parameter file is absent')
  call get_command_argument(1, paramfile)
  tag = ''
  tsim_max = 0d0
  limit_memory = 0d0
  do i = 2, nargs - 1, 2
     call get_command_argument(i, arg)
     select case (trim(adjustl(arg)))
     case ('-n')
        call get_command_argument(i + 1, tag)
     case ('-t')
        call get_command_argument(i + 1, arg)
        read(arg, *) tsim_max
     case ('-b')
        call get_command_argument(i + 1, arg)
        read(arg, *) limit_memory
     end select
  end do
  call system_clock(clock0, clock_rate)
  tsim_begin = 0.0

  ! ------------------------------------------------------------ keys (settings.f90:112-297)
  write(*, '(A)') ' #initialization'
  call ini_open(trim(paramfile), ok)
  if (.not. ok) then
     write(*, *) 'Ini_Open: Error opening file '//trim(paramfile)
     stop 1
  end if
  dry_run = ini_logical('dry_run')
  rootdir = ini_string('rootdir')
  if (len_trim(rootdir) == 0) then
     rootdir = '/'
  else if (rootdir(len_trim(rootdir):len_trim(rootdir)) /= '/') then
     rootdir = trim(rootdir)//'/'
  end if
  tmprn = ini_string('rootname')
  rootname = trim(rootdir)//trim(tmprn)
  if (tag /= '') then
     tmprn = trim(tmprn)//'_'//trim(tag)
     rootname = trim(rootname)//'_'//trim(tag)
  end if
  write(*, '(A)') ' rootname = '//trim(rootname)
  tmpdir = ini_string('tmpdir')
  if (tmpdir /= '') then
     if (tmpdir(len_trim(tmpdir):len_trim(tmpdir)) /= '/') tmpdir = trim(tmpdir)//'/'
  else
     tmpdir = rootdir
  end if
  ss = ini_string('ss')
  ! seed of the string fits' generator: two integers, or the clock when empty
  ! (init_rand_gen, subroutines.f90:61-74)
  if (ss /= '') then
     read(ss, *, iostat=info) seed(1:2)
     if (info /= 0) call die('error: cannot read seed; we need two integer*4')
  else
     call system_clock(count=clk)
     clk = mod(clk, 2147483648_8*2147483648_8)
     seed(2) = int(clk/2147483648_8, c_int32_t)
     seed(1) = int(clk - 2147483648_8*seed(2), c_int32_t)
  end if
  ngrid = ini_int('ngrid')
  if (mod(ngrid, 2_c_int64_t) /= 0) call die('error: ngrid should be even')
  n3 = ngrid**3
  n_chk = ini_int('n_chk')
  if (.not. n_chk >= 1) call die('error: n_chk should be 1 or larger')
  num_bin = ini_int('num_bin', 50)
  initial_loop = ini_logical('initial_loop', .false.)
  if (initial_loop) then
     b = ini_double('loop_radius')
     b = ini_double('loop_velocity')
     b = ini_double('loop_velocity_in')
  end if
  id_strings = ini_logical('id_strings')
  group_strings = .false.
  out_strings = .false.
  if (id_strings) then
     group_strings = ini_logical('group_strings')
     out_strings = ini_logical('out_strings')
  end if
  id_walls = ini_logical('id_walls')
  group_walls = .false.
  out_walls = .false.
  if (id_walls) then
     group_walls = ini_logical('group_walls')
     out_walls = ini_logical('out_walls')
  end if
  spec_est = ini_logical('spec_est')
  grav_spec = ini_logical('grav_spec')
  n_id = 0
  n_mod_eps = 0
  if (id_strings .or. id_walls .or. spec_est .or. grav_spec) then
     n_id = ini_int('n_id')
     if (spec_est) n_mod_eps = ini_int('n_mod_eps')
  end if
  degrade_spec = 0
  if (spec_est .or. grav_spec) then
     allocate(tau_id(0:n_chk))
     degrade_spec = ini_int('degrade_spec')
     if (degrade_spec < 0) call die('error: degrade_spec must be >= 0')
  end if
  if (spec_est) then
     if (n_mod_eps < 1) call die('error: n_mod_eps must be positive')
     ! crho and delta crho per snapshot (spectrum.f90:38-39), for get_epsilon at the end; a
     ! snapshot spectrum_estimation did not run for in this process (before a resume) keeps 0
     allocate(crho_spec(2, 0:n_chk))
     crho_spec = 0d0
  end if
  ! memory estimate (settings.f90:241-260), one task
  b = byte_dc*2
  if (spec_est) then
     b = b + 2*byte_dc/2**(degrade_spec*3)
     b = b + max(byte_dc/2**(degrade_spec*3), 1d0)
  else if (id_strings .or. id_walls) then
     b = b + 1d0
  end if
  b = b*dble(n3)/1024/1024/1024
  write(*, '(A,F12.4)') ' approx required total memory (GB) (without margin) = ', b
  if (b >= 2*hard_limit_cdata .and. .not. dry_run) call die('error: to large cdata; increase the number of tasks')
  b = b + 2*byte_dc*2*dble(ngrid*ngrid)/1024/1024/1024
  if (b >= limit_memory .and. .not. dry_run) call die('error: not enough memory')
  file_resume = trim(tmpdir)//trim(tmprn)//'_param.chk'
  sigma = ini_double('sigma')
  g_rel = ini_double('g_rel')
  kappa = ini_double('kappa')
  tau_ini = ini_double('tau_ini')
  tau_fin = ini_double('tau_fin')
  eps_dtau = ini_double('eps_dtau')
  boxsize = ini_double('boxsize')
  ! keys of this build
  prec = ini_string('precision')
  num_bin_grav = ini_int('num_bin_grav', 25)
  gw_fix_swap = ini_logical('gw_fix_swap', .false.)
  chk_gw = ini_logical('chk_gw', .false.)
  ic_seed = ini_int('ic_seed', 1234)
  ic_kmax = ini_double('ic_kmax', 1d0)
  ic_file = ini_string('ic_file')

  keys = axg_keys(sigma, g_rel, kappa, tau_fin, boxsize, ngrid)
  call axg_check(axg_derive(keys, prm, zeta, width_cr, thickness_cr), c_null_ptr, 'axg_derive')
  prm%precision = AXG_FP64
  if (trim(prec) == 'fp32c') prm%precision = AXG_FP32C
  prm%grav_spec = merge(1, 0, grav_spec)
  prm%initial_loop = merge(1, 0, initial_loop)
  prm%gw_fix_swap = merge(1, 0, gw_fix_swap)
  prm%num_bin_grav = num_bin_grav
  prm%num_bin = num_bin
  prm%id_strings = merge(1, 0, id_strings)
  prm%id_walls = merge(1, 0, id_walls)
  prm%spec_est = merge(1, 0, spec_est)
  prm%device = ini_int('device', 0)
  prm%eps_dtau = eps_dtau
  prm%dtau_fixed = ini_double('dtau_fixed', 0d0)
  prm%potential_kappa0 = ini_double('potential_kappa0', 0d0)
  prm%degrade_spec = degrade_spec
  ! the reference's build option VEL (Makefile:11) as a key: 1 = velocity_string_YY
  prm%string_vel = ini_int('string_vel', 0)
  call ini_close()
  write(*, '(A,I0)') ' ngrid = ', ngrid
  write(*, '(A,ES22.14)') ' sigma_cr = ', prm%sigma_cr
  write(*, '(A,ES22.14)') ' delta_lattice (in units of tau_cr) = ', prm%dspa
  if (dry_run) then
     write(*, '(A)') ' *** dry run just for checking parameters ***'
     stop
  end if

  call axg_check(axg_create(prm, h), c_null_ptr, 'axg_create')
  ! init_rand_gen (subroutines.f90:49-95): one MCG59 stream seeded by ss, from its first draw at
  ! every start (a resume re-seeds it too); the thermal state, the string fits and the
  ! covariance draw from it in the reference's order
  call axg_check(axg_set_stream(h, seed(1), seed(2), 0_c_int64_t), h, 'axg_set_stream')

  ! ------------------------------------------------------------ initial field (initial.f90:21-78)
  call set_initial_condition(chk)

  ! ------------------------------------------------------------ evolution (evolve.f90:26-152)
  tau0 = max(tau_ini, 1d0)
  dtau_chk = (tau_fin - tau0)/n_chk
  if (tau_ini + dtau_chk < 1d0) call die(' error: too large n_chk; tau at idx=1 is smaller than tau_at_pt')
  if (spec_est .or. grav_spec) then
     tau_id(0) = tau_ini
     do idx = 1, n_chk
        tau_id(idx) = tau0 + dtau_chk*idx
     end do
  end if
  if (chk) then
     call read_chk(tau_end, idx, flag_defects)
  else
     tau_end = tau_ini
     idx = 0
     flag_defects = .true.
     tsim_now = wall_hours()
     if ((tsim_now - tsim_begin)*(1 + 1d0/max(idx, 1)) > tsim_max) call write_chk(tau_end, idx, flag_defects)
  end if
  if (idx == 0) then
     open(300, file=trim(rootname)//'_string_wall_parameters.txt', position='rewind', iostat=info)
     open(400, file=trim(rootname)//'_crho.txt', position='rewind', iostat=info)
  else
     open(300, file=trim(rootname)//'_string_wall_parameters.txt', position='append', iostat=info)
     open(400, file=trim(rootname)//'_crho.txt', position='append', iostat=info)
  end if
  if (info /= 0) call die('error: cannot open file string_wall_parameters.txt')
  if (idx + n_id > n_chk .and. flag_defects) then
     call defects_and_spectrum(idx, tau_end)
     flag_defects = .false.
     call write_chk(tau_end, idx, flag_defects)
  end if
  ! the step kernels' launch configuration for this lattice, timed on the live fields (every
  ! configuration gives the same results)
  if (grav_spec) call axg_check(axg_tune(h, 3_c_int32_t, tune_code, c_null_ptr, 0_c_int32_t), h, 'axg_tune')
  do while (idx < n_chk)
     tau_start = tau_end
     idx = idx + 1
     tau_end = tau0 + dtau_chk*idx
     write(*, '(A,I3)') ' #idx = ', idx
     write(*, '(A,ES22.14,A,ES22.14)') '  tau_start = ', tau_start, ', tau_end = ', tau_end
     call axg_check(axg_advance(h, tau_start, tau_end, nsteps), h, 'axg_advance')
     write(*, '(A,I0)') '   iterations: ', nsteps
     tsim_now = wall_hours()
     if ((tsim_now - tsim_begin)*(1 + 1d0/idx) > tsim_max) then
        flag_defects = .true.
        call write_chk(tau_end, idx, flag_defects)
     end if
     if (idx + n_id > n_chk) then
        call defects_and_spectrum(idx, tau_end)
        if (tsim_now - tsim_begin > tsim_max) then
           flag_defects = .false.
           call write_chk(tau_end, idx, flag_defects)
        end if
     end if
  end do
  flag_defects = .false.
  call write_chk(tau_end, idx, flag_defects)
  close(300)
  close(400)
  if (spec_est) call get_epsilon()   ! main.f90:58 fin_spectrum
  call axg_check(axg_destroy(h), c_null_ptr, 'axg_destroy')
  write(*, '(A)') ' ... done'

contains

  ! print and stop with a failure status (the reference's MpiStop aborts the run)
  subroutine die(msg)
    character(*), intent(in) :: msg
    write(*, '(A)') ' '//msg
    stop 1
  end subroutine die

  ! wall time since start, in hours (the reference's MKL second()/3600)
  real function wall_hours()
    integer(8) :: c
    call system_clock(c)
    wall_hours = real(dble(c - clock0)/dble(clock_rate)/3600d0)
  end function wall_hours

  function itoa(n) result(s)
    integer, intent(in) :: n
    character(len=32) :: s
    write(s, '(I0)') n
  end function itoa

  function chk_name(n) result(s)
    integer, intent(in) :: n
    character(len=1100) :: s
    s = trim(tmpdir)//trim(tmprn)//'_phi_'//trim(itoa(n))//'.chk'
  end function chk_name

  function grav_chk_name(n) result(s)
    integer, intent(in) :: n
    character(len=1100) :: s
    s = trim(tmpdir)//trim(tmprn)//'_grav_'//trim(itoa(n))//'.chk'
  end function grav_chk_name

  ! initial.f90:21-78: resume when the param file and the named chk file exist; otherwise a
  ! fresh field (ic_file, else the seeded flat spectrum) and an empty param file
  subroutine set_initial_condition(chk)
    logical, intent(out) :: chk
    integer :: u, ios, idx0
    double precision :: t
    complex(c_double_complex), allocatable :: phi(:), pi(:)
    chk = .false.
    inquire(file=trim(file_resume), exist=exists)
    if (exists) then
       open(newunit=u, file=trim(file_resume), action='read', status='old', iostat=ios)
       read(u, *, iostat=ios) idx0, t
       close(u)
       if (ios == 0 .and. idx0 >= 0 .and. idx0 <= n_chk) then
          inquire(file=trim(chk_name(idx0)), exist=exists)
          if (exists) then
             write(*, '(A)') ' *** chk files EXIST!; restart from previous run ***'
             chk = .true.
             return
          end if
       end if
    end if
    write(*, '(A)') ' *generating new initial configuration ...'
    if (ic_file /= '') then
       allocate(phi(n3), pi(n3))
       open(newunit=u, file=trim(ic_file), access='stream', form='unformatted', status='old', action='read', &
            iostat=ios)
       if (ios /= 0) call die('error: cannot open ic_file')
       read(u) phi, pi
       close(u)
       call axg_check(axg_set_fields(h, phi, pi, n3), h, 'axg_set_fields')
       deallocate(phi, pi)
    else if (.not. initial_loop) then
       ! initial.f90:62-66: the thermal state before the transition
       write(*, '(A)') ' #starting from thermal distribution before phase transition'
       call axg_check(axg_init_thermal(h, ic_seed, tau_ini), h, 'axg_init_thermal')
    else
       ! in place of get_initial_loop (a relaxed string loop, not built here): the seeded
       ! flat-spectrum field
       call axg_check(axg_init_flat(h, ic_seed, ic_kmax), h, 'axg_init_flat')
    end if
    open(newunit=u, file=trim(file_resume), action='write', status='replace', iostat=ios)
    close(u)
    write(*, '(A)') '  *param.chk is created'
  end subroutine set_initial_condition

  ! evolve.f90:401-437: the param file is rewritten only if it exists (status='old'); fields
  ! are written when flag_defects, then older chk files are deleted
  subroutine write_chk(tau, idx, flag_defects)
    double precision, intent(in) :: tau
    integer, intent(in) :: idx
    logical, intent(in) :: flag_defects
    integer :: u, ios, iu
    if (flag_defects) call io_chk(.false., idx)
    open(newunit=u, file=trim(file_resume), action='write', status='old', iostat=ios)
    if (ios == 0) then
       write(u, '(I12,1X,ES25.16E3,1X,L2)', iostat=ios) idx, tau, flag_defects
       close(u)
    end if
    if (flag_defects) then
       do iu = 0, idx - 1
          open(newunit=u, file=trim(chk_name(iu)), status='old', iostat=ios)
          if (ios == 0) close(u, status='delete')
          open(newunit=u, file=trim(grav_chk_name(iu)), status='old', iostat=ios)
          if (ios == 0) close(u, status='delete')
       end do
    end if
  end subroutine write_chk

  ! evolve.f90:439-453
  subroutine read_chk(tau, idx, flag_defects)
    double precision, intent(out) :: tau
    integer, intent(out) :: idx
    logical, intent(out) :: flag_defects
    integer :: u, ios
    open(newunit=u, file=trim(file_resume), action='read', status='old', iostat=ios)
    if (ios == 0) read(u, *, iostat=ios) idx, tau, flag_defects
    if (ios /= 0) call die('error: cannot read param.chk')
    close(u)
    call io_chk(.true., idx)
  end subroutine read_chk

  ! evolve.f90:455-546: raw native FP64 complex, phi[N^3] then pi[N^3], x fastest
  subroutine io_chk(rd, idx)
    logical, intent(in) :: rd
    integer, intent(in) :: idx
    complex(c_double_complex), allocatable :: phi(:), pi(:), a(:)
    integer :: u, ios
    integer(c_int64_t) :: na, ng
    allocate(phi(n3), pi(n3))
    if (rd) then
       open(newunit=u, file=trim(chk_name(idx)), access='stream', form='unformatted', status='old', &
            action='read', iostat=ios)
       if (ios /= 0) call die('error: MPI_FILE_OPEN failed')
       read(u) phi, pi
       close(u)
       call axg_check(axg_set_fields(h, phi, pi, n3), h, 'axg_set_fields')
    else
       call axg_check(axg_get_fields(h, phi, pi, n3), h, 'axg_get_fields')
       open(newunit=u, file=trim(chk_name(idx)), access='stream', form='unformatted', status='replace', &
            action='write', iostat=ios)
       if (ios /= 0) call die('error: MPI_FILE_OPEN failed')
       write(u) phi, pi
       close(u)
    end if
    deallocate(phi, pi)
    if (chk_gw .and. grav_spec) then
       ! the accumulators live on the GW grid, ngrid / 2^degrade_spec per axis
       ng = ngrid/2**degrade_spec
       na = 4*ng*ng*(ng/2 + 1)
       allocate(a(na))
       if (rd) then
          inquire(file=trim(grav_chk_name(idx)), exist=exists)
          if (exists) then
             open(newunit=u, file=trim(grav_chk_name(idx)), access='stream', form='unformatted', status='old', &
                  action='read')
             read(u) a
             close(u)
             call axg_check(axg_set_gw_state(h, a, na), h, 'axg_set_gw_state')
          end if
       else
          call axg_check(axg_get_gw_state(h, a, na), h, 'axg_get_gw_state')
          open(newunit=u, file=trim(grav_chk_name(idx)), access='stream', form='unformatted', status='replace', &
               action='write')
          write(u) a
          close(u)
       end if
       deallocate(a)
    end if
  end subroutine io_chk

  ! evolve.f90:154-235 (DIM = 3) with id_defects' outputs, spectrum_estimation and
  ! grav_spec_estimation writing their files
  subroutine defects_and_spectrum(idx, tau)
    integer, intent(in) :: idx
    double precision, intent(in) :: tau
    type(axg_defect_out) :: dc
    type(axg_string_out) :: so
    type(axg_wall_out) :: wo
    type(axg_group_out) :: go
    double precision :: xi(3), largea(3), vel(3), sw(11), crho(3), rho, xi0, xi1, gam
    double precision, allocatable :: q(:), pw(:), nu(:), de(:), cv(:), nm(:)
    integer(c_int32_t), allocatable :: ml(:)
    integer(c_int64_t) :: nmask, k
    integer :: u, iq
    double precision :: freq
    xi = 0d0
    largea = 0d0
    vel = 0d0
    if (id_strings .or. id_walls .or. spec_est) then
       call axg_check(axg_defects(h, dc), h, 'axg_defects')
       write(*, '(A,3I12)') '  MpiRank    0 has found string, wall and mask points: ', dc%n_string, dc%n_wall, dc%n_mask
       if (spec_est) then
          allocate(ml(3))
          call axg_check(axg_get_mask(h, ml, 0_c_int64_t, nmask), h, 'axg_get_mask')
          deallocate(ml)
          allocate(ml(3*max(nmask, 1_c_int64_t)))
          call axg_check(axg_get_mask(h, ml, nmask, nmask), h, 'axg_get_mask')
          open(newunit=u, file=trim(rootname)//'_mask_'//trim(itoa(idx))//'-0.dat', action='write', status='replace')
          do k = 1, nmask
             write(u, '(3I6)') ml(3*k-2), ml(3*k-1), ml(3*k)
          end do
          close(u)
          deallocate(ml)
       end if
    end if
    ! string length and velocities (add_string_length, defects_3d.f90:276-279; without
    ! grouping the three lengths coincide); the generator's stream moves on with idx as the
    ! reference's does between calls.  Wall areas: add_wall_area (280-283).  With grouping
    ! (226-227) the second and third entries keep only the long strings / large walls.
    if (id_strings) then
       call axg_check(axg_strings(h, seed(1), seed(2) + int(idx, c_int32_t), so), h, 'axg_strings')
       xi(1:3) = so%length
       vel(1:3) = so%vel
       if (group_strings) then
          call axg_check(axg_group_strings(h, width_cr/prm%dspa/tau, tau/prm%dspa, go), h, 'axg_group_strings')
          write(*, '(A,F14.6)') ' size of string width in lattice units:', width_cr/prm%dspa/tau
          write(*, '(A,F14.6)') ' size of string separation in lattice units:', tau/prm%dspa
          write(*, '(A,F14.6)') ' threshold for string grouping in lattice units:', go%th_dist
          write(*, '(A,I0)') '  strings grouped into networks: ', go%n_group
          xi(1:3) = go%sum
       end if
       if (out_strings) call write_defects(0, idx, int(so%n_string, c_int64_t), group_strings)
       if (so%n_vel == 0) write(*, '(A)') ' velocity cannot be computed because there are no strings'
    end if
    if (id_walls) then
       call axg_check(axg_walls(h, wo), h, 'axg_walls')
       largea(1:3) = wo%area
       if (group_walls) then
          call axg_check(axg_group_walls(h, thickness_cr/prm%dspa/tau, tau/prm%dspa, go), h, 'axg_group_walls')
          write(*, '(A,F14.6)') ' size of wall width in lattice units:', thickness_cr/prm%dspa/tau
          write(*, '(A,F14.6)') ' size of wall separation in lattice units:', tau/prm%dspa
          write(*, '(A,F14.6)') ' threshold for wall grouping in lattice units:', go%th_dist
          write(*, '(A,I0)') '  walls grouped into networks: ', go%n_group
          largea(1:3) = go%sum
       end if
       if (out_walls) call write_defects(1, idx, int(wo%n_wall, c_int64_t), group_walls)
    end if
    xi(1:3) = (tau/prm%dspa)**2*xi(1:3)/dble(n3)/4
    largea(1:3) = (tau/prm%dspa)*largea(1:3)/dble(n3)/2
    if (id_strings) then  ! evolve.f90:192-196
       write(*, '(A,F14.6,3E20.10e3)') '  tau, string parameter: ', tau, xi(1:3)
       write(*, '(A,F14.6,3E20.10e3)') '  tau, string velocity parameter: ', tau, vel(1:3)
    end if
    if (id_walls) write(*, '(A,F14.6,3E20.10e3)') '  tau, wall parameter: ', tau, largea(1:3)
    sw = [tau, tau**2*zeta, xi(1:3), largea(1:3), vel(1:3)]
    crho = 0d0
    if (spec_est) then
       allocate(q(num_bin), pw(num_bin), cv(num_bin), nm(num_bin))
       call axg_check(axg_field_spectrum(h, tau_id(idx), q, pw, cv, nm, rho, num_bin), h, 'axg_field_spectrum')
       open(newunit=u, file=trim(rootname)//'_spectrum_'//trim(itoa(idx))//'.dat', action='write')
       do iq = 1, num_bin
          write(u, '(4E20.10e3)') q(iq), pw(iq), cv(iq)/nm(iq), nm(iq)
       end do
       close(u)
       call covariance(idx, rho)
       ! analyze_energy (evolve.f90:216-235)
       xi0 = sw(5)
       xi1 = sw(3) - sw(5)
       gam = sw(11)
       crho(1) = xi0*twopi_t*log(0.5d0*tau**2/sqrt(xi0)/width_cr)*4*gam
       crho(2) = xi1*twopi_t*log(0.5d0*tau**2/sqrt(xi0)/width_cr)*4*gam
       crho(3) = rho
       deallocate(q, pw, cv, nm)
    end if
    if (grav_spec) then
       allocate(q(num_bin_grav), pw(num_bin_grav), nu(num_bin_grav), de(num_bin_grav))
       call axg_check(axg_gw_spectrum(h, q, pw, nu, de, num_bin_grav), h, 'axg_gw_spectrum')
       freq = mstar_in_hz/prm%tau_cr/prm%a_now/twopi_t
       open(newunit=u, file=trim(rootname)//'_gravspec_'//trim(itoa(idx))//'.dat', action='write')
       do iq = 1, num_bin_grav
          write(u, '(5E20.10e3)') tau_id(idx), q(iq), q(iq)*freq, pw(iq), de(iq)
       end do
       close(u)
       deallocate(q, pw, nu, de)
    end if
    write(300, '(I12,11E20.10e3)') idx, sw(1:11)
    write(400, '(I12,3E20.10e3)') idx, crho(1:3)
    flush(300)
    flush(400)
  end subroutine defects_and_spectrum

  ! out_strings / out_walls (defects_3d.f90:232-260): one line per string, pos(3) dir(3) vel(3)
  ! gof sl '(3F10.3,8F8.3,2I4)', or per wall, pos(3) dir(3) gof wa '(3F10.3,5F8.3,2I4)', then the
  ! local and global group numbers (equal on one GPU; 0 without grouping).  The MPI build's
  ! concatenated file name _string_<idx>.dat / _wall_<idx>.dat is kept.  The reference writes
  ! the wall lines' group numbers from the wrong index (arr_wall(ns)); here each wall has its own.
  subroutine write_defects(kind, idx, n, grouped)
    integer, intent(in) :: kind, idx
    integer(c_int64_t), intent(in) :: n
    logical, intent(in) :: grouped
    double precision, allocatable :: rec(:)
    integer(c_int32_t), allocatable :: ids(:)
    integer(c_int64_t) :: cnt, k, w
    integer :: u
    w = merge(11_c_int64_t, 8_c_int64_t, kind == 0)
    allocate(rec(w*max(n, 1_c_int64_t)), ids(max(n, 1_c_int64_t)))
    ids = 0
    if (n > 0) then
       if (kind == 0) then
          call axg_check(axg_get_strings(h, rec, n, cnt), h, 'axg_get_strings')
       else
          call axg_check(axg_get_walls(h, rec, n, cnt), h, 'axg_get_walls')
       end if
       if (grouped) call axg_check(axg_get_groups(h, int(kind, c_int32_t), ids, n, cnt), h, 'axg_get_groups')
    end if
    if (kind == 0) then
       open(newunit=u, file=trim(rootname)//'_string_'//trim(itoa(idx))//'.dat', action='write', &
            status='replace', iostat=info)
       if (info /= 0) call die(' error: cannot create string.dat')
       do k = 1, n
          write(u, '(3F10.3,8F8.3,2I4)') rec(w*(k-1)+1:w*(k-1)+6), rec(w*(k-1)+9:w*(k-1)+11), &
               rec(w*(k-1)+7:w*(k-1)+8), ids(k), ids(k)
       end do
    else
       open(newunit=u, file=trim(rootname)//'_wall_'//trim(itoa(idx))//'.dat', action='write', &
            status='replace', iostat=info)
       if (info /= 0) call die(' error: cannot create wall.dat')
       do k = 1, n
          write(u, '(3F10.3,5F8.3,2I4)') rec(w*(k-1)+1:w*(k-1)+8), ids(k), ids(k)
       end do
    end if
    close(u)
    deallocate(rec, ids)
  end subroutine write_defects

  ! The rest of spectrum_estimation (spectrum.f90:290-309): estimate_covmat with min_mc = 10,
  ! max_mc = 100, tol = 0.1 on the GPU (its generator keyed by ss and idx), the covariance
  ! file '(2I12,E20.10e3)' (pairs j <= i), and compute_crho (316-342) for delta crho.
  subroutine covariance(idx, rho)
    integer, intent(in) :: idx
    double precision, intent(in) :: rho
    double precision, allocatable :: cov(:)
    integer(c_int32_t) :: iters
    integer(c_int64_t) :: ncov, key
    integer :: u, i, j, ij
    double precision :: drho, qc(num_bin), wb(num_bin)
    ncov = int(num_bin, c_int64_t)*(num_bin + 1)/2
    allocate(cov(ncov))
    write(*, '(A)') ' *estimating covariance matrix using MC (Gaussianity assumed)'
    key = ior(ishft(int(seed(1), c_int64_t), 32), iand(int(seed(2), c_int64_t) + idx, int(z'FFFFFFFF', c_int64_t)))
    call axg_check(axg_field_covariance(h, key, 10_c_int32_t, 100_c_int32_t, 1d-1, cov, ncov, iters), h, &
         'axg_field_covariance')
    if (iters > 0) then
       write(*, '(A,I0,A)') '  MC converged (', iters, ' realisations)'
    else
       write(*, '(A)') '  WARNING: MC failed to converge'
    end if
    open(newunit=u, file=trim(rootname)//'_covariance_'//trim(itoa(idx))//'.dat', action='write', iostat=info)
    if (info /= 0) call die('cannot open covariance.dat')
    do i = 1, num_bin
       do j = 1, i
          ij = i*(i - 1)/2 + j
          write(u, '(2I12,E20.10e3)') i, j, cov(ij)
       end do
    end do
    close(u)
    ! compute_crho's error (spectrum.f90:316-342): crho = sum_b P_b a_b / (2 pi^2) with
    ! a_b = qmid_b^2 dq_b, so delta crho = sqrt(a^T C a) / (2 pi^2)
    call spec_bins(qc, wb)
    drho = sqrt(packed_form(cov, qc*qc*wb, qc*qc*wb))/(2*pi_t*pi_t)
    crho_spec(1, idx) = rho
    crho_spec(2, idx) = drho
    write(*, '(A,3E16.4)') '  tau, crho, delta crho: ', tau_id(idx), crho_spec(1:2, idx)
    deallocate(cov)
  end subroutine covariance

  ! Centres and widths of the axion spectrum's bins.  The edges are those of init_spectrum
  ! (spectrum.f90:58-61): running sums of dq = pi/dspa_spec/num_bin from 0, dspa_spec = dspa
  ! 2^degrade_spec (spectrum.f90:46).
  subroutine spec_bins(qc, wb)
    double precision, intent(out) :: qc(num_bin), wb(num_bin)
    double precision :: lo, hi, step
    integer :: b
    step = pi_t/(prm%dspa*2**degrade_spec)/num_bin
    hi = 0d0
    do b = 1, num_bin
       lo = hi
       hi = hi + step
       qc(b) = (lo + hi)/2
       wb(b) = hi - lo
    end do
  end subroutine spec_bins

  ! x^T C y for the symmetric num_bin x num_bin matrix C stored packed-upper (the covariance
  ! files' order: entry (i, j <= i) at i(i-1)/2 + j).
  double precision function packed_form(cp, x, y)
    double precision, intent(in) :: cp(:), x(num_bin), y(num_bin)
    integer :: r, c, k
    packed_form = 0d0
    k = 0
    do r = 1, num_bin
       do c = 1, r
          k = k + 1
          if (c == r) then
             packed_form = packed_form + cp(k)*x(r)*y(r)
          else
             packed_form = packed_form + cp(k)*(x(r)*y(c) + x(c)*y(r))
          end if
       end do
    end do
  end function packed_form

  ! get_epsilon (spectrum.f90:956-1040), at the end of the run.  The last n_id snapshots' power
  ! (column 2 of _spectrum_<id>.dat) and covariance (column 3 of _covariance_<id>.dat) are read
  ! back.  For snapshots a and b = a + n_mod_eps, at tau = (tau_a + tau_b)/2 and f = tau/(4 pi):
  !   slope     eps  = f sum(dP q^2) / sum(dP q),  dP = P_b - P_a, q = the bin centres;
  !   its error deps = f sqrt(w^T (C_a + C_b) w) / |sum(dP q)|,  w = q^2 - eps q / f
  !   (eps's linearisation in the two independent spectra);
  ! with the change of crho and its error in quadrature, as '(2I12,7E20.10e3)' rows of _eps.txt.
  subroutine get_epsilon()
    integer :: a, b, k, u, first, ncov
    double precision, allocatable :: pw(:, :), cv(:, :)
    double precision :: qc(num_bin), wb(num_bin), dpw(num_bin), w(num_bin)
    double precision :: tmid, f, slope, eps, deps, c1, c2
    write(*, *)
    write(*, '(A)') ' #epsilon'
    first = n_chk - n_id + 1
    ncov = num_bin*(num_bin + 1)/2
    allocate(pw(num_bin, first:n_chk), cv(ncov, first:n_chk))
    do a = first, n_chk
       open(newunit=u, file=trim(rootname)//'_spectrum_'//trim(itoa(a))//'.dat', action='read', status='old', &
            iostat=info)
       if (info /= 0) call die(' error: spectrum file '//trim(itoa(a))//' is missing')
       do k = 1, num_bin
          read(u, *, iostat=info) c1, pw(k, a)
          if (info /= 0) call die(' error: short spectrum file '//trim(itoa(a)))
       end do
       close(u)
       open(newunit=u, file=trim(rootname)//'_covariance_'//trim(itoa(a))//'.dat', action='read', status='old', &
            iostat=info)
       if (info /= 0) call die(' error: covariance file '//trim(itoa(a))//' is missing')
       do k = 1, ncov
          read(u, *, iostat=info) c1, c2, cv(k, a)
          if (info /= 0) call die(' error: short covariance file '//trim(itoa(a)))
       end do
       close(u)
    end do
    call spec_bins(qc, wb)
    open(newunit=u, file=trim(rootname)//'_eps.txt', action='write', status='replace', iostat=info)
    if (info /= 0) call die(' error: cannot create the eps file')
    do a = first, n_chk - n_mod_eps, n_mod_eps
       b = a + n_mod_eps
       tmid = (tau_id(b) + tau_id(a))/2
       f = tmid/(2*twopi_t)
       dpw = pw(:, b) - pw(:, a)
       slope = sum(dpw*qc)
       eps = f*sum(dpw*qc*qc)/slope
       w = qc*qc - eps*qc/f
       deps = f*sqrt(packed_form(cv(:, a) + cv(:, b), w, w))/abs(slope)
       write(u, '(2I12,7E20.10e3)') a, b, tau_id(a), tau_id(b), tmid, eps, deps, &
            crho_spec(1, b) - crho_spec(1, a), hypot(crho_spec(2, b), crho_spec(2, a))
    end do
    close(u)
    deallocate(pw, cv)
  end subroutine get_epsilon

end program axion_3d
