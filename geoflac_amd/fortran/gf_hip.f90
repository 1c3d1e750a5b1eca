! gf_hip.f90 -- ISO_C_BINDING interface of libgfhip.so (include/gf_abi.h).
!
! The derived types mirror the C structs field by field (same order, same kinds), so
! a Fortran caller fills them from its own module variables and passes them by
! reference.  Host arrays are passed as raw addresses of the reference's column-major
! arrays (src/arrays.f90:9-53); the library never keeps Fortran descriptors.
module gf_hip
use iso_c_binding
implicit none

integer, parameter :: GF_MAXPH = 20, GF_MAXINH = 100
integer(c_int32_t), parameter :: GF_PREC_FP64 = 0, GF_PREC_MIXED = 1, GF_PREC_MIXED_STORAGE = 2
integer(c_int32_t), parameter :: GF_ALL = int(z'3FFFFF', c_int32_t)
integer(c_int32_t), parameter :: GF_MARKERS = int(z'400000', c_int32_t)
integer(c_int32_t), parameter :: GF_PHASE_RATIO = int(z'200000', c_int32_t), GF_IPHASE = int(z'100000', c_int32_t)
integer(c_int32_t), parameter :: GF_APS = int(z'4000', c_int32_t)
integer(c_int32_t), parameter :: GF_STRESSBC = int(z'1000000', c_int32_t)
integer(c_int32_t), parameter :: GF_TOPO = int(z'2000000', c_int32_t)
integer, parameter :: GF_OK = 0, GF_STOP_MAXWELL = 22, GF_STOP_NEGATIVE_AREA = 40

type, bind(C) :: gf_params_t
    integer(c_int32_t) :: nx, nz, nphase
    integer(c_int32_t) :: mix_strain, mix_stress, ndim, movegrid
    integer(c_int32_t) :: itherm, istress_therm, ishearh, itemp_bc, itype_melting
    integer(c_int32_t) :: nystressbc, nydrsides, nyhydro, i_prestress
    integer(c_int32_t) :: idt_scale, i_rey, ivis_present, iphsub
    integer(c_int32_t) :: ifreq_rmasses, ifreq_imasses, ifreq_visc, ifreq_avgsr
    integer(c_int32_t) :: inhom
    integer(c_int32_t) :: irheol(GF_MAXPH)
    integer(c_int32_t) :: ix1(GF_MAXINH), ix2(GF_MAXINH), iy1(GF_MAXINH), iy2(GF_MAXINH), igeom(GF_MAXINH)
    integer(c_int32_t) :: flexure_base
    integer(c_int32_t) :: nelem_serp, nopbmax
    real(c_double) :: t_top, t_bot, bot_bc, g, demf, ten_off, tau_heal, v_min, v_max
    real(c_double) :: frac, fracm, strain_inert, dt_scale, xReyn, vbc, rxbo, rzbo
    real(c_double) :: pisos, drosub, fmagma_max, weaken_ratio_plastic, weaken_ratio_viscous, rho_magma
    real(c_double) :: den(GF_MAXPH), alfa(GF_MAXPH), beta(GF_MAXPH), pln(GF_MAXPH), acoef(GF_MAXPH)
    real(c_double) :: eactiv(GF_MAXPH), vactiv(GF_MAXPH), rl(GF_MAXPH), rm(GF_MAXPH)
    real(c_double) :: plstrain1(GF_MAXPH), plstrain2(GF_MAXPH), fric1(GF_MAXPH), fric2(GF_MAXPH)
    real(c_double) :: cohesion1(GF_MAXPH), cohesion2(GF_MAXPH), dilat1(GF_MAXPH), dilat2(GF_MAXPH)
    real(c_double) :: conduct(GF_MAXPH), cp(GF_MAXPH)
    real(c_double) :: topo_kappa
    real(c_double) :: erodibility(GF_MAXPH)
end type

type, bind(C) :: gf_host_arrays_t
    type(c_ptr) :: cord, vel, temp, temp0, force, amass, rmass, bc
    type(c_ptr) :: ncod
    type(c_ptr) :: stress0, area, dvol, strain, se2sr
    type(c_ptr) :: aps, visn, e2sr, sshrheat, shrheat, source
    type(c_ptr) :: iphase
    type(c_ptr) :: phase_ratio
    type(c_ptr) :: nopbou, bcstress
    type(c_ptr) :: coh_scale
    type(c_ptr) :: dtopo, dhacc, dhacc_correction
end type

type, bind(C) :: gf_marker_arrays_t
    integer(c_int32_t) :: nmarkers, capacity
    type(c_ptr) :: mark_a1, mark_a2, mark_ntriag, mark_phase, mark_dead, mark_id_elem, nmark_elem
    type(c_ptr) :: mark_age
end type

type, bind(C) :: gf_clock_t
    integer(c_int32_t) :: nloop, nsrate
    real(c_double) :: time, dt, dtavg, dt_elastic, dt_maxwell
end type

type, bind(C) :: gf_output_opts_t
    integer(c_int32_t) :: io_vel, io_srII, io_eII, io_aps, io_sII, io_sxx, io_szz, io_sxz, io_pres, io_temp, io_phase
    integer(c_int32_t) :: io_visc, io_unused, io_density, io_src, io_diss, io_forc, io_hfl, io_topo
    integer(c_int32_t) :: lastout
    real(c_double) :: hs
    type(c_ptr) :: xoriginal, zoriginal, dhacc, extr_acc, dhacc_correction, q_init, vel_flex_old
    type(c_ptr) :: fmagma, fmagma2, dike_backlog, dike_marker_vol, mor_marker_vol, mark_age
end type

interface
    integer(c_int) function gf_create(p, precision, ctx) bind(C, name='gf_create')
        import :: c_int, c_int32_t, c_ptr, gf_params_t
        type(gf_params_t), intent(in) :: p
        integer(c_int32_t), value :: precision
        type(c_ptr), intent(out) :: ctx
    end function
    subroutine gf_destroy(ctx) bind(C, name='gf_destroy')
        import :: c_ptr
        type(c_ptr), value :: ctx
    end subroutine
    type(c_ptr) function gf_last_error(ctx) bind(C, name='gf_last_error')
        import :: c_ptr
        type(c_ptr), value :: ctx
    end function
    integer(c_int) function gf_error_site(ctx, j, i) bind(C, name='gf_error_site')
        import :: c_int, c_int32_t, c_ptr
        type(c_ptr), value :: ctx
        integer(c_int32_t), intent(out) :: j, i
    end function
    integer(c_int) function gf_bind_host(ctx, h) bind(C, name='gf_bind_host')
        import :: c_int, c_ptr, gf_host_arrays_t
        type(c_ptr), value :: ctx
        type(gf_host_arrays_t), intent(in) :: h
    end function
    integer(c_int) function gf_bind_markers(ctx, m) bind(C, name='gf_bind_markers')
        import :: c_int, c_ptr, gf_marker_arrays_t
        type(c_ptr), value :: ctx
        type(gf_marker_arrays_t), intent(in) :: m
    end function
    integer(c_int) function gf_marker_count(ctx, n) bind(C, name='gf_marker_count')
        import :: c_int, c_int32_t, c_ptr
        type(c_ptr), value :: ctx
        integer(c_int32_t), intent(out) :: n
    end function
    integer(c_int) function gf_itest_mesh(ctx, mode_rem, dx_rem, angle_rem, ratio_mantle_mzone, need, anglemin, &
                                          val) bind(C, name='gf_itest_mesh')
        import :: c_int, c_int32_t, c_double, c_ptr
        type(c_ptr), value :: ctx
        integer(c_int32_t), value :: mode_rem
        real(c_double), value :: dx_rem, angle_rem, ratio_mantle_mzone
        integer(c_int32_t), intent(out) :: need
        real(c_double), intent(out) :: anglemin, val
    end function
    integer(c_int) function gf_upload(ctx, mask) bind(C, name='gf_upload')
        import :: c_int, c_int32_t, c_ptr
        type(c_ptr), value :: ctx
        integer(c_int32_t), value :: mask
    end function
    integer(c_int) function gf_download(ctx, mask) bind(C, name='gf_download')
        import :: c_int, c_int32_t, c_ptr
        type(c_ptr), value :: ctx
        integer(c_int32_t), value :: mask
    end function
    integer(c_int) function gf_set_clock(ctx, c) bind(C, name='gf_set_clock')
        import :: c_int, c_ptr, gf_clock_t
        type(c_ptr), value :: ctx
        type(gf_clock_t), intent(in) :: c
    end function
    integer(c_int) function gf_get_clock(ctx, c) bind(C, name='gf_get_clock')
        import :: c_int, c_ptr, gf_clock_t
        type(c_ptr), value :: ctx
        type(gf_clock_t), intent(out) :: c
    end function
    integer(c_int) function gf_step(ctx, nsteps, c) bind(C, name='gf_step')
        import :: c_int, c_int32_t, c_ptr, gf_clock_t
        type(c_ptr), value :: ctx
        integer(c_int32_t), value :: nsteps
        type(gf_clock_t), intent(inout) :: c
    end function
    integer(c_int) function gf_flac(ctx, c) bind(C, name='gf_flac')
        import :: c_int, c_ptr, gf_clock_t
        type(c_ptr), value :: ctx
        type(gf_clock_t), intent(inout) :: c
    end function
    integer(c_int) function gf_outflac(ctx, c, dir, o) bind(C, name='gf_outflac')
        import :: c_int, c_ptr, c_char, gf_clock_t, gf_output_opts_t
        type(c_ptr), value :: ctx
        type(gf_clock_t), intent(in) :: c
        character(kind=c_char), dimension(*), intent(in) :: dir
        type(gf_output_opts_t), intent(in) :: o
    end function
    integer(c_int) function gf_saveflac(ctx, c, dir, o) bind(C, name='gf_saveflac')
        import :: c_int, c_ptr, c_char, gf_clock_t, gf_output_opts_t
        type(c_ptr), value :: ctx
        type(gf_clock_t), intent(in) :: c
        character(kind=c_char), dimension(*), intent(in) :: dir
        type(gf_output_opts_t), intent(in) :: o
    end function
    integer(c_int) function gf_sync(ctx) bind(C, name='gf_sync')
        import :: c_int, c_ptr
        type(c_ptr), value :: ctx
    end function
end interface

contains

    ! Address of a contiguous double precision array (no copy: assumed size).
    function gf_addr8(a) result(p)
        real(c_double), target, intent(in) :: a(*)
        type(c_ptr) :: p
        p = c_loc(a)
    end function

    function gf_addr4(a) result(p)
        integer(c_int32_t), target, intent(in) :: a(*)
        type(c_ptr) :: p
        p = c_loc(a)
    end function

    function gf_addr1(a) result(p)
        integer(c_int8_t), target, intent(in) :: a(*)
        type(c_ptr) :: p
        p = c_loc(a)
    end function

    ! The C string of gf_last_error as a Fortran string.
    function gf_message(ctx) result(msg)
        type(c_ptr), intent(in) :: ctx
        character(len=512) :: msg
        character(kind=c_char), pointer :: s(:)
        type(c_ptr) :: cs
        integer :: k
        msg = ' '
        cs = gf_last_error(ctx)
        if (.not. c_associated(cs)) return
        call c_f_pointer(cs, s, [512])
        do k = 1, 512
            if (s(k) == c_null_char) exit
            msg(k:k) = s(k)
        enddo
    end function

end module gf_hip
