! The reference-side binding with the reference's own field arrays, as a runnable check.
!
! The reference holds phi and pi in data1/data2, shaped (ngrid, ngrid, local_0+2) with one halo
! plane on each side: interior planes min_0 = 2 .. max_0 = local_0+1 (initial.f90:240-285; on
! one rank local_0 = ngrid).  The C-ABI takes the N^3 interior, so the binding passes the
! sections data1(:,:,2:ngrid+1) and data2(:,:,2:ngrid+1), and refreshes the halo planes
! (update_margin, evolve.f90:341-393) after reading the fields back.
!
! The check builds such arrays over one flat buffer (as c_f_pointer does over the FFTW
! allocation), fills the halo planes with a sentinel and puts a domain wall between interior
! planes ngrid/2 and ngrid/2+1 (phase -1.6 below, +1.6 above: the largest gap of a cell that
! straddles it does not cross +-pi).  The GPU must flag exactly the cells of 1-based z = ngrid/2
! and, through the periodic wrap, z = ngrid as walls: a one-plane shift of the fields moves
! them.  Reading back into another halo-padded array must return the interior bit for bit and
! leave its halo planes alone.  Prints "halo binding ok" and stops 0, or stops 1.
program halo_binding_check
  use, intrinsic :: iso_c_binding
  use axgrav_capi
  implicit none
  integer(c_int64_t), parameter :: ngrid = 16
  integer(c_int64_t) :: n3, ix, iy, iz
  complex(c_double_complex), allocatable, target :: buf1(:), buf2(:), buf3(:), buf4(:)
  complex(c_double_complex), pointer :: data1(:,:,:), data2(:,:,:), data3(:,:,:), data4(:,:,:)
  integer(c_int8_t), allocatable :: flags(:,:,:)
  complex(c_double_complex), parameter :: sentinel = (7d0, -7d0)
  type(axg_params) :: prm
  type(axg_defect_out) :: dc
  type(c_ptr) :: h
  real(c_double) :: ph, twopi_t
  logical :: ok, want

  n3 = ngrid**3
  allocate(buf1(ngrid*ngrid*(ngrid+2)), buf2(ngrid*ngrid*(ngrid+2)))
  allocate(buf3(ngrid*ngrid*(ngrid+2)), buf4(ngrid*ngrid*(ngrid+2)))
  call c_f_pointer(c_loc(buf1), data1, [ngrid, ngrid, ngrid+2])
  call c_f_pointer(c_loc(buf2), data2, [ngrid, ngrid, ngrid+2])
  call c_f_pointer(c_loc(buf3), data3, [ngrid, ngrid, ngrid+2])
  call c_f_pointer(c_loc(buf4), data4, [ngrid, ngrid, ngrid+2])
  data1 = sentinel
  data2 = sentinel
  data3 = sentinel
  data4 = sentinel
  do iz = 1, ngrid
     ph = merge(1.6d0, -1.6d0, iz > ngrid/2)
     do iy = 1, ngrid
        do ix = 1, ngrid
           data1(ix, iy, iz+1) = cmplx(cos(ph), sin(ph), c_double_complex)
           data2(ix, iy, iz+1) = cmplx(real(ix, c_double), real(iz + 100*iy, c_double), c_double_complex)
        end do
     end do
  end do

  twopi_t = 2*3.14159265359d0
  prm%ngrid = ngrid
  prm%precision = AXG_FP64
  prm%grav_spec = 0
  prm%initial_loop = 1
  prm%id_strings = 1
  prm%id_walls = 1
  prm%spec_est = 1
  prm%dspa = 0.25d0
  prm%dwav = twopi_t/prm%dspa/ngrid
  prm%dv_spa = prm%dspa**3
  prm%dv_wav = (prm%dwav/twopi_t)**3
  prm%sigma_cr = sqrt(0.5d0)
  prm%tau_cr = 1
  prm%a_now = 1
  prm%kappa = 1d-30
  call axg_check(axg_create(prm, h), c_null_ptr, 'axg_create')
  call axg_check(axg_set_fields(h, data1(:,:,2:ngrid+1), data2(:,:,2:ngrid+1), n3), h, 'axg_set_fields')

  allocate(flags(ngrid, ngrid, ngrid))
  call axg_check(axg_defects(h, dc), h, 'axg_defects')
  call axg_check(axg_get_flags(h, flags, n3), h, 'axg_get_flags')
  ok = dc%n_wall == 2*ngrid*ngrid
  do iz = 1, ngrid
     want = iz == ngrid/2 .or. iz == ngrid
     do iy = 1, ngrid
        do ix = 1, ngrid
           ok = ok .and. (iand(int(flags(ix, iy, iz)), 2) /= 0) .eqv. want
        end do
     end do
  end do
  if (.not. ok) then
     write(*, '(A,I0)') ' error: wall cells not on z = ngrid/2 and ngrid; walls flagged: ', dc%n_wall
     stop 1
  end if

  call axg_check(axg_get_fields(h, data3(:,:,2:ngrid+1), data4(:,:,2:ngrid+1), n3), h, 'axg_get_fields')
  ok = all(data3(:,:,2:ngrid+1) == data1(:,:,2:ngrid+1)) .and. all(data4(:,:,2:ngrid+1) == data2(:,:,2:ngrid+1))
  ok = ok .and. all(data3(:,:,1) == sentinel) .and. all(data3(:,:,ngrid+2) == sentinel)
  ok = ok .and. all(data4(:,:,1) == sentinel) .and. all(data4(:,:,ngrid+2) == sentinel)
  if (.not. ok) then
     write(*, '(A)') ' error: the interior did not come back in place, or a halo plane was written'
     stop 1
  end if
  ! update_margin on one rank: the halo planes are the periodic neighbours of the interior
  data3(:,:,1) = data3(:,:,ngrid+1)
  data3(:,:,ngrid+2) = data3(:,:,2)
  call axg_check(axg_destroy(h), c_null_ptr, 'axg_destroy')
  write(*, '(A)') ' halo binding ok'
end program halo_binding_check
