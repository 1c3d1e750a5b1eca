! Test harness for the compiled GeoFLAC reference (test infrastructure only).
!
! This program is NOT part of the product.  It is linked against the
! reference's own subroutines, compiled from the reference's source files
! where they lie (see oracle/Makefile), and replaces only the reference's
! main program (src/par.f90:3-137) with a loop that runs an exact number of
! steps and dumps the complete in-memory state, so that golden vectors hold
! every array the explicit step reads or writes.
!
! The step loop follows src/par.f90:81-91 (flac; nloop += 1; time += dt)
! without periodic output.  With GF_REM=1 it also runs the driver's remeshing
! block after every ntest_rem steps (src/par.f90:94-110: itest_mesh, then
! bar2euler, remesh, lpeuler2bar, marker2elem when needed).  Compiled with
! -DGF_DROPIN the same harness drives the HIP drop-in `flac`
! (geoflac_amd/fortran/flac_hip.f90), syncs device state to the Fortran arrays
! before each dump and before the remeshing test, and back to the device after a
! remeshing, as a driver using the drop-in does.
!
! usage: ref_harness <inputfile> <outdir> <nsteps> [dump_step ...]
program gf_ref_harness
use arrays
use params
use marker_data
implicit none
character(len=200) :: inputfile
character(len=512) :: outdir, arg
integer :: nsteps, nd, k, istep, narg, nrem
logical :: do_rem
integer, allocatable :: dumps(:)

narg = command_argument_count()
if (narg < 3) then
    write(*,*) 'usage: ref_harness inputfile outdir nsteps [dump_step ...]'
    stop 1
endif
call get_command_argument(1, inputfile)
call get_command_argument(2, outdir)
call get_command_argument(3, arg)
read(arg, *) nsteps
nd = narg - 3
allocate(dumps(max(nd, 1)))
dumps = -1
do k = 1, nd
    call get_command_argument(3 + k, arg)
    read(arg, *) dumps(k)
enddo

open(unit=333, file='output.asc')

call read_params(inputfile)
call allocate_arrays(nz, nx, nphase)
call allocate_markers(nz, nx)
! fields the reference never initialises on a new start but the step reads
force = 0.d0
dike_released = 0.d0
call setflac
call dump_params(outdir)
! every dump directory is made here, before the first step: once the drop-in has
! initialised the GPU this process starts no other program
do k = 1, nd
    write(arg, '(A,A,I7.7)') trim(outdir), '/step_', dumps(k)
    call execute_command_line('mkdir -p '//trim(arg))
enddo
if (any(dumps(1:nd) == 0)) call dump_state(outdir, 0)

call get_environment_variable('GF_REM', arg)
do_rem = trim(arg) == '1'
nrem = 0
do istep = 1, nsteps
    call flac
    nloop = nloop + 1
    time = time + dt
    if (do_rem .and. mod(istep, ntest_rem) == 0) call remesh_block
    if (any(dumps(1:nd) == istep)) then
#ifdef GF_DROPIN
        call flac_hip_sync_host
#endif
        call dump_state(outdir, istep)
    endif
enddo
! GF_REF_WRITE=1: the reference's own writers on the final state, into the working
! directory (outflac's frames, saveflac's restart files)
call get_environment_variable('GF_REF_WRITE', arg)
if (trim(arg) == '1') then
#ifdef GF_DROPIN
    call flac_hip_sync_host
#endif
    call outflac
    call saveflac
endif
if (do_rem) write(*, '(A,I6)') 'remeshings ', nrem
#ifdef GF_DROPIN
! GF_HIP_WRITE=1: the drop-in's writers from the device state instead
call get_environment_variable('GF_HIP_WRITE', arg)
if (trim(arg) == '1') then
    call flac_hip_outflac
    call flac_hip_saveflac
endif
call flac_hip_finalize
#endif
close(333)

contains

! src/par.f90:94-110
subroutine remesh_block
    integer :: need_remeshing
    if (ny_rem /= 1 .or. itherm == 2) return
#ifdef GF_DROPIN
    ! the test on the device; the state comes back only for a remeshing
    call flac_hip_itest_mesh(need_remeshing)
    if (need_remeshing == 0) return
    call flac_hip_sync_host
#else
    call itest_mesh(need_remeshing)
    if (need_remeshing == 0) return
#endif
    nrem = nrem + 1
    call bar2euler
    call remesh
    call lpeuler2bar
    call marker2elem
    dhacc_correction(:) = 0.d0
#ifdef GF_DROPIN
    call flac_hip_sync_device
#endif
end subroutine

subroutine wr8(dir, name, a, n)
    character(len=*), intent(in) :: dir, name
    integer, intent(in) :: n
    double precision, intent(in) :: a(n)
    integer :: u
    open(newunit=u, file=trim(dir)//'/'//name//'.f64', access='stream', form='unformatted', status='replace')
    write(u) a
    close(u)
end subroutine

subroutine wr4(dir, name, a, n)
    character(len=*), intent(in) :: dir, name
    integer, intent(in) :: n
    integer, intent(in) :: a(n)
    integer :: u
    open(newunit=u, file=trim(dir)//'/'//name//'.i32', access='stream', form='unformatted', status='replace')
    write(u) a
    close(u)
end subroutine

subroutine dump_state(base, step)
    character(len=*), intent(in) :: base
    integer, intent(in) :: step
    character(len=600) :: dir
    integer :: u, ne, nn
    write(dir, '(A,A,I7.7)') trim(base), '/step_', step
    ne = (nz-1)*(nx-1)
    nn = nz*nx
    call wr8(dir, 'cord', cord, nn*2)
    call wr8(dir, 'vel', vel, nn*2)
    call wr8(dir, 'temp', temp, nn)
    call wr8(dir, 'temp0', temp0, nn)
    call wr8(dir, 'force', force, nn*2)
    call wr8(dir, 'amass', amass, nn)
    call wr8(dir, 'rmass', rmass, nn)
    call wr8(dir, 'bc', bc, nn*2)
    call wr4(dir, 'ncod', ncod, nn*2)
    call wr8(dir, 'stress0', stress0, ne*16)
    call wr8(dir, 'area', area, ne*4)
    call wr8(dir, 'dvol', dvol, ne*4)
    call wr8(dir, 'strain', strain, ne*3)
    call wr8(dir, 'se2sr', se2sr, ne*3)
    call wr8(dir, 'aps', aps, ne)
    call wr8(dir, 'visn', visn, ne)
    call wr8(dir, 'e2sr', e2sr, ne)
    call wr8(dir, 'sshrheat', sshrheat, ne)
    call wr8(dir, 'shrheat', shrheat, ne)
    call wr8(dir, 'source', source, ne)
    call wr4(dir, 'iphase', iphase, ne)
    call wr8(dir, 'phase_ratio', phase_ratio, ne*nphase)
    ! markers (module marker_data, src/marker_data.f90:1-20): the first nmarkers entries
    call wr4(dir, 'mark_phase', int(mark_phase(1:nmarkers)), nmarkers)
    call wr4(dir, 'nmark_elem', nmark_elem, ne)
    ! surface processes (diff_topo, src/fl_move.f90:138-315): the topography change of the
    ! step and its accumulations, and the markers resurface added or removed
    if (topo_kappa > 0.d0) then
        call wr8(dir, 'dtopo', dtopo, nx)
        call wr8(dir, 'dhacc', dhacc, nx-1)
        call wr8(dir, 'dhacc_correction', dhacc_correction, nx)
        call wr8(dir, 'extr_acc', extr_acc, nx-1)
        call wr8(dir, 'mark_age', mark_age, nmarkers)
        if (step > 0) then
            call wr8(dir, 'mark_a1', mark_a1, nmarkers)
            call wr8(dir, 'mark_a2', mark_a2, nmarkers)
            call wr4(dir, 'mark_ntriag', mark_ntriag, nmarkers)
            call wr4(dir, 'mark_dead', int(mark_dead(1:nmarkers)), nmarkers)
            call wr4(dir, 'mark_id_elem', mark_id_elem, max_markers_per_elem*ne)
        endif
    endif
    if (step == 0) then
        ! stress boundary segments (init_bc, src/init_bc.f90:226-280)
        call wr4(dir, 'nopbou', nopbou, size(nopbou))
        call wr8(dir, 'bcstress', bcstress, size(bcstress))
        call wr8(dir, 'mark_a1', mark_a1, nmarkers)
        call wr8(dir, 'mark_a2', mark_a2, nmarkers)
        call wr4(dir, 'mark_ntriag', mark_ntriag, nmarkers)
        call wr4(dir, 'mark_dead', int(mark_dead(1:nmarkers)), nmarkers)
        call wr4(dir, 'mark_id_elem', mark_id_elem, max_markers_per_elem*ne)
    endif
    open(newunit=u, file=trim(dir)//'/scalars.txt', status='replace')
    write(u, '(A,I12)') 'nloop ', nloop
    write(u, '(A,I12)') 'nsrate ', nsrate
    write(u, '(A,1X,ES25.17E3)') 'time ', time
    write(u, '(A,1X,ES25.17E3)') 'dt ', dt
    write(u, '(A,1X,ES25.17E3)') 'dtavg ', dtavg
    write(u, '(A,1X,ES25.17E3)') 'dt_elastic ', dt_elastic
    write(u, '(A,1X,ES25.17E3)') 'dt_maxwell ', dt_maxwell
    write(u, '(A,1X,ES25.17E3)') 'frac ', frac
    write(u, '(A,I12)') 'nmarkers ', nmarkers
    close(u)
end subroutine

subroutine dump_params(base)
    character(len=*), intent(in) :: base
    integer :: u, k
    call execute_command_line('mkdir -p '//trim(base))
    open(newunit=u, file=trim(base)//'/params.txt', status='replace')
    write(u, '(A,3I12)') 'nx_nz_nphase ', nx, nz, nphase
    write(u, '(A,99I12)') 'flags ', mix_strain, mix_stress, ndim, movegrid, itherm, istress_therm, &
        ishearh, itemp_bc, itype_melting, nystressbc, nydrsides, nyhydro, i_prestress, idt_scale, &
        i_rey, ivis_present, iphsub, ifreq_rmasses, ifreq_imasses, ifreq_visc, ifreq_avgsr, inhom
    write(u, '(A,99(1X,ES25.17E3))') 'reals ', t_top, t_bot, bot_bc, g, demf, ten_off, tau_heal, v_min, &
        v_max, frac, fracm, strain_inert, dt_scale, xReyn, vbc, rxbo, rzbo, pisos, drosub, &
        fmagma_max, weaken_ratio_plastic, weaken_ratio_viscous, rho_magma, extra_pres
    write(u, '(A,99I12)') 'flags2 ', nelem_serp, nopbmax
    do k = 1, nphase
        write(u, '(A,I4,I4,99(1X,ES25.17E3))') 'phase ', k, irheol(k), den(k), alfa(k), beta(k), pln(k), &
            acoef(k), eactiv(k), vactiv(k), rl(k), rm(k), plstrain1(k), plstrain2(k), fric1(k), &
            fric2(k), cohesion1(k), cohesion2(k), dilat1(k), dilat2(k), conduct(k), cp(k)
    enddo
    write(u, '(A,99(1X,ES25.17E3))') 'topo ', topo_kappa, (erodibility(k), k = 1, nphase)
    do k = 1, inhom
        write(u, '(A,6I8,1X,ES25.17E3)') 'inhom ', ix1(k), ix2(k), iy1(k), iy2(k), inphase(k), igeom(k), xinitaps(k)
    enddo
    close(u)
end subroutine

end program gf_ref_harness
