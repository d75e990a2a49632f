! ref_fortran_harness.f90 -- test infrastructure only (never part of the
! product, never run on the GPU host).  A driver for the reference's own
! Fortran-lineage subroutines, compiled where they lie by oracle/Makefile:
!   init_cpml, CPML_E4, CPML_H4        ver0040/cpu_cpml4.f90
!   media_coeff, e_field_cpml4,
!   h_field_cpml4                      ver0040/cpu_eh_field_cpml42.f90
!   read_source_3d                     ver0040/cpu_waveform.f90
! on the compile-time grid of ver0040/cpu_parameter_wafe.f90 (module
! const_para).  It calls them in the order of the reference main loop
! (ver0040/cpu_main.f90:96-99, 104-140) on inputs read from a file, and writes
! the results for tests/golden/make_fortran_fixtures.py.  It computes nothing
! itself.
!
! usage: fortran_step <input file> <output file>   (run in a scratch directory:
! the reference subroutines write their own text files into the cwd)
!
! input (stream, native endianness): nsteps (int32), six 1-based source
! coordinates x_source, y_source, z_source, x_source2, y_source2, z_source2
! (int32), sig(nx,ny,nz), myu(nx,ny,nz), then Ex, Ey, Ez, Hx, Hy, Hz as their
! real parts followed by their imaginary parts (real(8), x fastest).
! output: dt, cmax, omega0, MU0, c1, c2, dx (real(8)); per axis x, y, z the
! profiles esig, msig, be, bh, ce, ch, ked, khd; the coefficient volumes
! ca_x..z, cb_x..z, da_x..z, db_x..z; the six fields (real, imaginary); the
! twelve psi volumes psi_Ezx1, psi_Eyx1, psi_Exy1, psi_Ezy1, psi_Eyz1,
! psi_Exz1, psi_Hyx1, psi_Hzx1, psi_Hzy1, psi_Hxy1, psi_Hxz1, psi_Hyz1 (real,
! imaginary).
program ref_fortran_harness
    use const_para
    implicit none

    complex(kind(0d0)), allocatable :: f(:,:,:,:)
    complex(kind(0d0)), allocatable :: Je(:), Jh(:)
    real(8), allocatable :: re(:,:,:), im(:,:,:)
    integer :: nsteps, istep, q, src(6)
    real(8) :: t
    character(len=1024) :: fin, fout

    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    allocate(f(nx, ny, nz, 6), re(nx, ny, nz), im(nx, ny, nz), Je(nstep), Jh(nstep))

    open(10, file=trim(fin), access='stream', form='unformatted', status='old')
    read(10) nsteps
    read(10) src
    read(10) sig
    read(10) myu
    do q = 1, 6
        read(10) re
        read(10) im
        f(:, :, :, q) = cmplx(re, im, kind(0d0))
    enddo
    close(10)
    if (nsteps < 0 .or. nsteps > nstep) stop 2

    x_source = src(1)
    y_source = src(2)
    z_source = src(3)
    x_source2 = src(4)
    y_source2 = src(5)
    z_source2 = src(6)
    psi_Ezx1 = (0.0d0, 0.0d0)
    psi_Eyx1 = (0.0d0, 0.0d0)
    psi_Exy1 = (0.0d0, 0.0d0)
    psi_Ezy1 = (0.0d0, 0.0d0)
    psi_Eyz1 = (0.0d0, 0.0d0)
    psi_Exz1 = (0.0d0, 0.0d0)
    psi_Hyx1 = (0.0d0, 0.0d0)
    psi_Hzx1 = (0.0d0, 0.0d0)
    psi_Hzy1 = (0.0d0, 0.0d0)
    psi_Hxy1 = (0.0d0, 0.0d0)
    psi_Hxz1 = (0.0d0, 0.0d0)
    psi_Hyz1 = (0.0d0, 0.0d0)
    grad = 0.0d0

    call init_cpml
    call media_coeff
    t = 0.0d0
    do istep = 1, nsteps
        call read_source_3d(istep, t, f(:, :, :, 6), Je, Jh)
        call e_field_cpml4(f(:, :, :, 1), f(:, :, :, 2), f(:, :, :, 3), &
                           f(:, :, :, 4), f(:, :, :, 5), f(:, :, :, 6))
        call CPML_E4(f(:, :, :, 1), f(:, :, :, 2), f(:, :, :, 3), &
                     f(:, :, :, 4), f(:, :, :, 5), f(:, :, :, 6))
        t = t + dt * 0.5d0
        call h_field_cpml4(f(:, :, :, 1), f(:, :, :, 2), f(:, :, :, 3), &
                           f(:, :, :, 4), f(:, :, :, 5), f(:, :, :, 6))
        call CPML_H4(f(:, :, :, 1), f(:, :, :, 2), f(:, :, :, 3), &
                     f(:, :, :, 4), f(:, :, :, 5), f(:, :, :, 6))
        t = t + dt * 0.5d0
    enddo

    open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(11) dt, cmax, omega0, MU0, c1, c2, dx
    write(11) esig_x, msig_x, be_x, bh_x, ce_x, ch_x, kedx, khdx
    write(11) esig_y, msig_y, be_y, bh_y, ce_y, ch_y, kedy, khdy
    write(11) esig_z, msig_z, be_z, bh_z, ce_z, ch_z, kedz, khdz
    write(11) ca_x, ca_y, ca_z, cb_x, cb_y, cb_z, da_x, da_y, da_z, db_x, db_y, db_z
    do q = 1, 6
        write(11) real(f(:, :, :, q), kind(0d0)), aimag(f(:, :, :, q))
    enddo
    write(11) real(psi_Ezx1, kind(0d0)), aimag(psi_Ezx1), real(psi_Eyx1, kind(0d0)), aimag(psi_Eyx1)
    write(11) real(psi_Exy1, kind(0d0)), aimag(psi_Exy1), real(psi_Ezy1, kind(0d0)), aimag(psi_Ezy1)
    write(11) real(psi_Eyz1, kind(0d0)), aimag(psi_Eyz1), real(psi_Exz1, kind(0d0)), aimag(psi_Exz1)
    write(11) real(psi_Hyx1, kind(0d0)), aimag(psi_Hyx1), real(psi_Hzx1, kind(0d0)), aimag(psi_Hzx1)
    write(11) real(psi_Hzy1, kind(0d0)), aimag(psi_Hzy1), real(psi_Hxy1, kind(0d0)), aimag(psi_Hxy1)
    write(11) real(psi_Hxz1, kind(0d0)), aimag(psi_Hxz1), real(psi_Hyz1, kind(0d0)), aimag(psi_Hyz1)
    close(11)
end program ref_fortran_harness
