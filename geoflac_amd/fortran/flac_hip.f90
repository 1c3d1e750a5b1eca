! flac_hip.f90 -- drop-in replacement of subroutine flac (src/flac.f90:5-91).
!
! Same name, no arguments, same place in the driver's loop (src/par.f90:83): each call
! runs one explicit step (fl_therm, change_phase, fl_rheol, fl_node, fl_move, and
! rmasses / dt_mass at their frequencies) on the GPU through libgfhip.so.  The markers
! of module marker_data go to the device with the arrays; change_phase updates their
! phases there.  The first call copies the
! parameters of module params and every array of module arrays to the device; after
! that the device state is authoritative and the Fortran arrays are refreshed only by
! flac_hip_sync_host, which the driver calls where it reads them (output, restart
! dumps, remeshing: src/par.f90:94-127).  After the host has changed arrays itself
! (remeshing, src/remesh.f90) it calls flac_hip_sync_device.  nloop and time stay
! owned by the driver; nsrate, dt, dtavg, dt_elastic and dt_maxwell come back from
! every call, as the reference's flac leaves them in module params.
!
! Errors: the reference's STOP codes (22, src/rh_maxwell.f90:29; 40,
! src/fl_move.f90:125) are raised the same way, after SysMsg; ABI errors (>= 1000)
! stop with their code and the library's message.
module flac_hip_state
use iso_c_binding
implicit none
type(c_ptr), save :: gf_ctx_handle = c_null_ptr
end module flac_hip_state

subroutine flac
use iso_c_binding
use gf_hip
use flac_hip_state
use params
implicit none
type(gf_clock_t) :: ck
integer(c_int) :: rc

if (.not. c_associated(gf_ctx_handle)) call flac_hip_init
ck%nloop = nloop
ck%nsrate = nsrate
ck%time = time
ck%dt = dt
ck%dtavg = dtavg
ck%dt_elastic = dt_elastic
ck%dt_maxwell = dt_maxwell
rc = gf_flac(gf_ctx_handle, ck)
if (rc /= GF_OK) call flac_hip_fail(rc)
nsrate = ck%nsrate
dt = ck%dt
dtavg = ck%dtavg
dt_elastic = ck%dt_elastic
dt_maxwell = ck%dt_maxwell
end subroutine flac

! Parameters and arrays to the device (the state setflac/rsflac left on the host).
subroutine flac_hip_init
use iso_c_binding
use gf_hip
use flac_hip_state
use arrays
use params
implicit none
type(gf_params_t) :: p
type(gf_host_arrays_t) :: h
type(gf_clock_t) :: ck
integer(c_int) :: rc
integer :: k, n

p%nx = nx; p%nz = nz; p%nphase = nphase
p%mix_strain = mix_strain; p%mix_stress = mix_stress; p%ndim = ndim; p%movegrid = movegrid
p%itherm = itherm; p%istress_therm = istress_therm; p%ishearh = ishearh; p%itemp_bc = itemp_bc
p%itype_melting = itype_melting; p%nystressbc = nystressbc; p%nydrsides = nydrsides
p%nyhydro = nyhydro; p%i_prestress = i_prestress; p%idt_scale = idt_scale; p%i_rey = i_rey
p%ivis_present = ivis_present; p%iphsub = iphsub; p%ifreq_rmasses = ifreq_rmasses
p%ifreq_imasses = ifreq_imasses; p%ifreq_visc = ifreq_visc; p%ifreq_avgsr = ifreq_avgsr
p%inhom = inhom
p%nelem_serp = nelem_serp
p%nopbmax = nopbmax
p%irheol = 0; p%ix1 = 0; p%ix2 = 0; p%iy1 = 0; p%iy2 = 0; p%igeom = 0
! the flexural base (src/fl_flexure.f90:15-24) is refused by gf_create
p%flexure_base = 0
do k = 1, nofbc
    if (nbc(k) == 200 .and. nofside(k) == 2) p%flexure_base = 1
enddo
p%topo_kappa = topo_kappa
n = min(maxph, GF_MAXPH)
p%irheol(1:n) = irheol(1:n)
n = min(maxinh, GF_MAXINH)
p%ix1(1:n) = ix1(1:n); p%ix2(1:n) = ix2(1:n); p%iy1(1:n) = iy1(1:n); p%iy2(1:n) = iy2(1:n)
p%igeom(1:n) = igeom(1:n)
p%t_top = t_top; p%t_bot = t_bot; p%bot_bc = bot_bc; p%g = g; p%demf = demf
p%ten_off = ten_off; p%tau_heal = tau_heal; p%v_min = v_min; p%v_max = v_max
p%frac = frac; p%fracm = fracm; p%strain_inert = strain_inert; p%dt_scale = dt_scale
p%xReyn = xReyn; p%vbc = vbc; p%rxbo = rxbo; p%rzbo = rzbo; p%pisos = pisos; p%drosub = drosub
p%fmagma_max = fmagma_max; p%weaken_ratio_plastic = weaken_ratio_plastic
p%weaken_ratio_viscous = weaken_ratio_viscous; p%rho_magma = rho_magma
do k = 1, GF_MAXPH
    if (k > maxph) exit
    p%den(k) = den(k); p%alfa(k) = alfa(k); p%beta(k) = beta(k); p%pln(k) = pln(k)
    p%acoef(k) = acoef(k); p%eactiv(k) = eactiv(k); p%vactiv(k) = vactiv(k)
    p%rl(k) = rl(k); p%rm(k) = rm(k); p%plstrain1(k) = plstrain1(k); p%plstrain2(k) = plstrain2(k)
    p%fric1(k) = fric1(k); p%fric2(k) = fric2(k); p%cohesion1(k) = cohesion1(k)
    p%cohesion2(k) = cohesion2(k); p%dilat1(k) = dilat1(k); p%dilat2(k) = dilat2(k)
    p%conduct(k) = conduct(k); p%cp(k) = cp(k); p%erodibility(k) = erodibility(k)
enddo

rc = gf_create(p, GF_PREC_FP64, gf_ctx_handle)
if (rc /= GF_OK) call flac_hip_fail(rc)

h%cord = gf_addr8(cord); h%vel = gf_addr8(vel); h%temp = gf_addr8(temp); h%temp0 = gf_addr8(temp0)
h%force = gf_addr8(force); h%amass = gf_addr8(amass); h%rmass = gf_addr8(rmass); h%bc = gf_addr8(bc)
h%ncod = gf_addr4(ncod)
h%stress0 = gf_addr8(stress0); h%area = gf_addr8(area); h%dvol = gf_addr8(dvol)
h%strain = gf_addr8(strain); h%se2sr = gf_addr8(se2sr)
h%aps = gf_addr8(aps); h%visn = gf_addr8(visn); h%e2sr = gf_addr8(e2sr)
h%sshrheat = gf_addr8(sshrheat); h%shrheat = gf_addr8(shrheat); h%source = gf_addr8(source)
h%iphase = gf_addr4(iphase)
h%phase_ratio = gf_addr8(phase_ratio)
h%nopbou = gf_addr4(nopbou)
h%bcstress = gf_addr8(bcstress)
h%coh_scale = c_null_ptr
h%dtopo = gf_addr8(dtopo); h%dhacc = gf_addr8(dhacc); h%dhacc_correction = gf_addr8(dhacc_correction)
rc = gf_bind_host(gf_ctx_handle, h)
if (rc /= GF_OK) call flac_hip_fail(rc)
call flac_hip_bind_markers
rc = gf_upload(gf_ctx_handle, ior(ior(ior(GF_ALL, GF_MARKERS), GF_STRESSBC), merge(GF_TOPO, 0_c_int32_t, topo_kappa > 0.d0)))
if (rc /= GF_OK) call flac_hip_fail(rc)
ck%nloop = nloop; ck%nsrate = nsrate; ck%time = time; ck%dt = dt; ck%dtavg = dtavg
ck%dt_elastic = dt_elastic; ck%dt_maxwell = dt_maxwell
rc = gf_set_clock(gf_ctx_handle, ck)
if (rc /= GF_OK) call flac_hip_fail(rc)
end subroutine flac_hip_init

! The marker arrays of module marker_data (their count changes with remeshing).
subroutine flac_hip_bind_markers
use iso_c_binding
use gf_hip
use flac_hip_state
use marker_data
use params
implicit none
type(gf_marker_arrays_t) :: m
integer(c_int) :: rc
m%nmarkers = nmarkers
m%capacity = max_markers   ! (src/marker_data.f90:27) resurface appends up to it
m%mark_a1 = gf_addr8(mark_a1); m%mark_a2 = gf_addr8(mark_a2)
m%mark_ntriag = gf_addr4(mark_ntriag)
m%mark_phase = gf_addr1(mark_phase); m%mark_dead = gf_addr1(mark_dead)
m%mark_id_elem = gf_addr4(mark_id_elem); m%nmark_elem = gf_addr4(nmark_elem)
m%mark_age = gf_addr8(mark_age)
rc = gf_bind_markers(gf_ctx_handle, m)
if (rc /= GF_OK) call flac_hip_fail(rc)
end subroutine flac_hip_bind_markers

! Device -> Fortran arrays (before outflac / saveflac / remeshing read them).  With
! surface processes the markers resurface added come back too, with their count.
subroutine flac_hip_sync_host
use iso_c_binding
use gf_hip
use flac_hip_state
use params
use marker_data
implicit none
integer(c_int) :: rc
integer(c_int32_t) :: n
if (.not. c_associated(gf_ctx_handle)) return
rc = gf_download(gf_ctx_handle, ior(ior(GF_ALL, GF_MARKERS), merge(GF_TOPO, 0_c_int32_t, topo_kappa > 0.d0)))
if (rc /= GF_OK) call flac_hip_fail(rc)
rc = gf_marker_count(gf_ctx_handle, n)
if (rc /= GF_OK) call flac_hip_fail(rc)
nmarkers = n
end subroutine flac_hip_sync_host

! The driver's remeshing test, itest_mesh (src/rem_test.f90:9-139), on the device's
! coordinates, with the reference's messages: a driver copies the state back
! (flac_hip_sync_host) only when it remeshes.
subroutine flac_hip_itest_mesh(need_remeshing)
use iso_c_binding
use gf_hip
use flac_hip_state
use params
implicit none
integer, intent(out) :: need_remeshing
integer(c_int) :: rc
integer(c_int32_t) :: need
real(c_double) :: anglemint, val
need_remeshing = 0
if (.not. c_associated(gf_ctx_handle)) return
rc = gf_itest_mesh(gf_ctx_handle, int(mode_rem, c_int32_t), dx_rem, angle_rem, ratio_mantle_mzone, need, &
                   anglemint, val)
if (rc /= GF_OK) call flac_hip_fail(rc)
need_remeshing = need
if (need_remeshing == 2) then
    if( dtout_screen .ne. 0 ) then
        print *, 'Remeshing due to shortening required: ', val
        write(333,*) 'Remeshing due to shortening required: ', val
    else
        call SysMsg('TEST_MESH: Remeshing due to shortening required')
    endif
    return
endif
if (need_remeshing == 3) then
    if( dtout_screen .ne. 0 ) then
        print *, 'Remeshing due to surface topo changes required: ', val
        write(333,*) 'Remeshing due to surface topo changes required: ', val
    else
        call SysMsg('TEST_MESH: Remeshing due to surface topo changes required')
    endif
    return
endif
if( dtout_screen .ne. 0 ) then
    write (6,'(A,F6.2,A,F10.6)') '        min.angle=',anglemint, '     dt(yr)=',dt/sec_year
    write (333,'(A,F6.2,A,F10.6)') '        min.angle=',anglemint, '     dt(yr)=',dt/sec_year
    flush (333)
endif
if (need_remeshing == 1) then
    if( dtout_screen .ne. 0 ) then
        print *, 'Remeshing due to angle required.'
        write(333,*) 'Remeshing due to angle required.'
    else
        call SysMsg('TEST_MESH: Remeshing due to angle required.')
    endif
endif
end subroutine flac_hip_itest_mesh

! Fortran arrays -> device (after the host changed them, e.g. remeshing).
subroutine flac_hip_sync_device
use iso_c_binding
use gf_hip
use flac_hip_state
use params
implicit none
integer(c_int) :: rc
type(gf_clock_t) :: ck
if (.not. c_associated(gf_ctx_handle)) return
call flac_hip_bind_markers
rc = gf_upload(gf_ctx_handle, ior(ior(GF_ALL, GF_MARKERS), merge(GF_TOPO, 0_c_int32_t, topo_kappa > 0.d0)))
if (rc /= GF_OK) call flac_hip_fail(rc)
ck%nloop = nloop; ck%nsrate = nsrate; ck%time = time; ck%dt = dt; ck%dtavg = dtavg
ck%dt_elastic = dt_elastic; ck%dt_maxwell = dt_maxwell
rc = gf_set_clock(gf_ctx_handle, ck)
if (rc /= GF_OK) call flac_hip_fail(rc)
end subroutine flac_hip_sync_device

! The reference's output and restart writers (src/outflac.f90, src/saveflac.f90) from
! the device state, into the working directory, without copying the state back:
! outflac's frames are evaluated on the device.  The arrays the explicit step does
! not change come from the Fortran modules.  Amin.rs, a list-directed scalar of
! setflac, is still written by the caller's own code.
subroutine flac_hip_output_opts(o)
use iso_c_binding
use gf_hip
use arrays
use params
use marker_data
implicit none
type(gf_output_opts_t), intent(out) :: o
o%io_vel = io_vel; o%io_srII = io_srII; o%io_eII = io_eII; o%io_aps = io_aps; o%io_sII = io_sII
o%io_sxx = io_sxx; o%io_szz = io_szz; o%io_sxz = io_sxz; o%io_pres = io_pres; o%io_temp = io_temp
o%io_phase = io_phase; o%io_visc = io_visc; o%io_unused = io_unused; o%io_density = io_density
o%io_src = io_src; o%io_diss = io_diss; o%io_forc = io_forc; o%io_hfl = io_hfl; o%io_topo = io_topo
o%lastout = lastout
o%hs = hs
o%xoriginal = gf_addr8(xoriginal); o%zoriginal = gf_addr8(zoriginal)
o%dhacc = gf_addr8(dhacc); o%extr_acc = gf_addr8(extr_acc); o%dhacc_correction = gf_addr8(dhacc_correction)
o%q_init = gf_addr8(q_init); o%vel_flex_old = gf_addr8(vel_flex_old)
o%fmagma = gf_addr8(fmagma); o%fmagma2 = gf_addr8(fmagma2)
o%dike_backlog = gf_addr8(dike_backlog); o%dike_marker_vol = gf_addr8(dike_marker_vol)
o%mor_marker_vol = gf_addr8(mor_marker_vol)
o%mark_age = gf_addr8(mark_age)
end subroutine flac_hip_output_opts

subroutine flac_hip_outflac
use iso_c_binding
use gf_hip
use flac_hip_state
use params
implicit none
type(gf_output_opts_t) :: o
type(gf_clock_t) :: ck
integer(c_int) :: rc
if (.not. c_associated(gf_ctx_handle)) call flac_hip_init
call flac_hip_output_opts(o)
ck%nloop = nloop; ck%nsrate = nsrate; ck%time = time; ck%dt = dt; ck%dtavg = dtavg
ck%dt_elastic = dt_elastic; ck%dt_maxwell = dt_maxwell
rc = gf_outflac(gf_ctx_handle, ck, '.'//c_null_char, o)
if (rc /= GF_OK) call flac_hip_fail(rc)
end subroutine flac_hip_outflac

subroutine flac_hip_saveflac
use iso_c_binding
use gf_hip
use flac_hip_state
use params
implicit none
type(gf_output_opts_t) :: o
type(gf_clock_t) :: ck
integer(c_int) :: rc
if (.not. c_associated(gf_ctx_handle)) call flac_hip_init
call flac_hip_output_opts(o)
ck%nloop = nloop; ck%nsrate = nsrate; ck%time = time; ck%dt = dt; ck%dtavg = dtavg
ck%dt_elastic = dt_elastic; ck%dt_maxwell = dt_maxwell
rc = gf_saveflac(gf_ctx_handle, ck, '.'//c_null_char, o)
if (rc /= GF_OK) call flac_hip_fail(rc)
end subroutine flac_hip_saveflac

subroutine flac_hip_finalize
use iso_c_binding
use gf_hip
use flac_hip_state
implicit none
if (.not. c_associated(gf_ctx_handle)) return
call gf_destroy(gf_ctx_handle)
gf_ctx_handle = c_null_ptr
end subroutine flac_hip_finalize

subroutine flac_hip_fail(rc)
use iso_c_binding
use gf_hip
use flac_hip_state
implicit none
integer(c_int), intent(in) :: rc
integer(c_int32_t) :: j, i
integer(c_int) :: r2
character(len=600) :: msg
j = 0; i = 0
if (c_associated(gf_ctx_handle)) r2 = gf_error_site(gf_ctx_handle, j, i)
if (rc == GF_STOP_MAXWELL) then
    write(msg, '(A,2I8)') 'Maxwell: time step!!! (element j, i)', j, i
    call SysMsg(trim(msg))
    stop 22
else if (rc == GF_STOP_NEGATIVE_AREA) then
    write(msg, '(A,2I8)') 'FL_MOVE: NEGATIVE AREA! (element j, i)', j, i
    call SysMsg(trim(msg))
    stop 40
endif
if (c_associated(gf_ctx_handle)) then
    msg = 'libgfhip: '//trim(gf_message(gf_ctx_handle))
else
    msg = 'libgfhip: context creation failed'
endif
call SysMsg(trim(msg))
write(*, '(A,I6)') trim(msg)//' code', rc
error stop 1
end subroutine flac_hip_fail
