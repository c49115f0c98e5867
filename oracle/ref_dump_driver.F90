! Test harness for the reference build in oracle/_ref (test infrastructure only;
! never shipped, never linked into the product).
!
! It drives the reference's own modules exactly the way the reference program
! does (src/main.F90:70-121: initParams -> setupVelocityGrid ->
! setupVirtualProcessGrid -> setupPhysicalGrid -> setupFlow -> allocateBuf, then
! the iStep loop of iterate / chkConverge / convergence exit), and in addition
! writes the solver state as raw little-endian binary so that the C oracle and
! the GPU path can be compared at full double precision (the reference's own
! outputs are ES15.6 text, src/output.F90:156-158).
!
! Usage (run in a directory holding para.in and the image):
!     piks_ref_dump OUTDIR [fSTEP | mSTEP | ffinal | mfinal ...]
! writes OUTDIR/setup_*.bin once; after iteration STEP it writes f, fw and the
! moments (fSTEP) or the moments only (mSTEP); ffinal/mfinal do the same for
! the state the loop ends with.  OUTDIR/conv.bin holds one record per
! chkConverge call: (int32 iStep, float64 mass, float64 permeability,
! float64 error).
program piks_ref_dump
    use parameters
    use flow
    use mpiParams
    use solver
    use physicalGrid
    use velocityGrid
    use output
    implicit none
    include "mpif.h"

    integer :: ierr, provided, nargs, a, ndump, u, laststep
    integer, allocatable :: dumpat(:), dumpfull(:)
    logical :: final_f, final_m
    character(len=512) :: outdir, arg

    call MPI_INIT_THREAD(MPI_THREAD_FUNNELED, provided, ierr)
    call MPI_COMM_SIZE(MPI_COMM_WORLD, nprocs, ierr)
    call MPI_COMM_RANK(MPI_COMM_WORLD, proc, ierr)

    nargs = command_argument_count()
    if (nargs < 1) then
        print *, 'usage: piks_ref_dump OUTDIR [fSTEP|mSTEP|ffinal|mfinal ...]'
        stop 1
    endif
    call get_command_argument(1, outdir)
    ndump = nargs - 1
    allocate(dumpat(max(ndump, 1)), dumpfull(max(ndump, 1)))
    dumpat = -1
    dumpfull = 0
    final_f = .false.
    final_m = .false.
    do a = 1, ndump
        call get_command_argument(a + 1, arg)
        if (trim(arg) == 'ffinal') then
            final_f = .true.
        else if (trim(arg) == 'mfinal') then
            final_m = .true.
        else
            read(arg(2:), *) dumpat(a)
            if (arg(1:1) == 'f') dumpfull(a) = 1
        endif
    enddo

    call initParams
    call setupVelocityGrid
    call setupVirtualProcessGrid
    call setupPhysicalGrid
    call setupFlow
    call allocateBuf

    call dump_setup()

    open(newunit=u, file=trim(outdir)//'/conv.bin', access='stream', &
         form='unformatted', status='replace')

    error = 1.d0
    laststep = 0
    do iStep = 1, maxStep
        call iterate
        laststep = iStep
        if (mod(iStep, chkConvergeStep) == 0) then
            call chkConverge
            write(u) iStep, mass, permeability, error
        endif
        do a = 1, ndump
            if (dumpat(a) == iStep) call dump_step(iStep, dumpfull(a) == 1)
        enddo
        if (error <= eps) then
            if (iStep < chkConvergeStep) then
                call chkConverge
                write(u) iStep, mass, permeability, error
            endif
            exit
        endif
    enddo
    close(u)
    if (final_f .or. final_m) call dump_step(-1, final_f)
    write(*,'(A,I8,A,ES25.17)') 'DUMP_DONE steps=', laststep, ' K=', permeability

    call MPI_FINALIZE(ierr)

contains

    subroutine wint(name, arr)
        character(len=*), intent(in) :: name
        integer, intent(in) :: arr(:)
        integer :: v
        open(newunit=v, file=trim(outdir)//'/'//name, access='stream', &
             form='unformatted', status='replace')
        write(v) arr
        close(v)
    end subroutine wint

    subroutine wdbl(name, arr)
        character(len=*), intent(in) :: name
        double precision, intent(in) :: arr(:)
        integer :: v
        open(newunit=v, file=trim(outdir)//'/'//name, access='stream', &
             form='unformatted', status='replace')
        write(v) arr
        close(v)
    end subroutine wdbl

    subroutine dump_setup()
        integer :: v
        call wint('setup_image.bin', reshape(image, (/ size(image) /)))
        call wint('setup_which_corner.bin', reshape(which_corner, (/ size(which_corner) /)))
        call wint('setup_vecWall.bin', vecWall)
        call wdbl('setup_extCoef.bin', reshape(extCoef, (/ size(extCoef) /)))
        call wint('setup_mapF.bin', mapF)
        call wdbl('setup_cx.bin', cx)
        call wdbl('setup_cy.bin', cy)
        call wdbl('setup_cz.bin', cz)
        call wdbl('setup_w.bin', w)
        call wint('setup_dir1.bin', reshape(dir1, (/ size(dir1) /)))
        call wint('setup_dir8.bin', reshape(dir8, (/ size(dir8) /)))
        call wdbl('setup_coef1.bin', reshape(coef1, (/ size(coef1) /)))
        call wdbl('setup_coef7.bin', reshape(coef7, (/ size(coef7) /)))
        open(newunit=v, file=trim(outdir)//'/setup_scalars.txt', status='replace')
        write(v,'(A,7I12)') 'grid ', Nx, Ny, Nz, Nxtotal, Nytotal, Nztotal, Ntotal
        write(v,'(A,12I8)') 'bounds ', xl, xu, yl, yu, zl, zu, xlg, xug, ylg, yug, zlg, zug
        write(v,'(A,3I12)') 'counts ', nWall, Nfluid, size(fw, 1)
        write(v,'(A,8I12)') 'nstencil ', Nstencil1, Nstencil2, Nstencil3, Nstencil4, &
                            Nstencil5, Nstencil6, Nstencil7, Nstencil8
        write(v,'(A,2I8)') 'nc ', Nc, Nc8
        write(v,'(A,ES26.17E3)') 'ds ', ds
        write(v,'(A,ES26.17E3)') 'DiffFlux ', DiffFlux
        write(v,'(A,ES26.17E3)') 'mu ', mu
        close(v)
    end subroutine dump_setup

    subroutine dump_step(s, full)
        integer, intent(in) :: s
        logical, intent(in) :: full
        character(len=32) :: tag
        if (s < 0) then
            tag = 'final'
        else
            write(tag, '(I0)') s
        endif
        if (full) then
            call wdbl('f_'//trim(tag)//'.bin', reshape(f, (/ size(f) /)))
            if (size(fw) > 0) call wdbl('fw_'//trim(tag)//'.bin', reshape(fw, (/ size(fw) /)))
        endif
        call wdbl('rho_'//trim(tag)//'.bin', Rho)
        call wdbl('ux_'//trim(tag)//'.bin', Ux)
        call wdbl('uy_'//trim(tag)//'.bin', Uy)
        call wdbl('uz_'//trim(tag)//'.bin', Uz)
    end subroutine dump_step

end program piks_ref_dump
