! piks3d_gpu -- drop-in replacement of the reference program (src/main.F90)
! running the DVM iteration on an MI355X through libpiks3d.so.
!
! Same surface as the reference: it reads para.in in the working directory
! (the five namelists of src/parameters.F90:78-82, same defaults; the mpiNml
! sizes are accepted and ignored because one GPU holds the whole domain) and
! the raw 0/1 image named there, runs the reference's loop (src/main.F90:95-121:
! iterate, chkConverge every chkConvergeStep, save every saveStep, stop when
! error <= eps), prints the per-check line and appends Results.dat like
! src/output.F90:156-159, prints the Walltime/Steps/K line of
! src/main.F90:126-127, and writes the flow field as VTI (+PVTI), Tecplot or
! legacy VTK in the formats of src/output.F90:166-360.
!
! One optional namelist is added, ignored by the reference:
!     &gpuNml precision = 'fp64' | 'fp32', device = 0, exact_quadrature = F,
!             exact_order = F, refineSwitch = 1.0d-6 /
! exact_order = T sums the moments and wall fluxes in the reference's order,
! so every value (and every output file) equals the reference's bit for bit.
! With precision = 'fp32' the library runs its FP64 defect correction (FP32
! correction iterations between FP64 residual iterations, tracking the FP64
! iterates; see piks_set_precision in include/piks3d.h); refineSwitch is
! read and ignored.
program piks3d_gpu
    use iso_c_binding
    use piks_c
    implicit none

    ! para.in, names as in the reference
    character(len=40) :: imageFileName
    integer :: Nx, Ny, Nz, wallExtOrder, fluidlayer
    double precision :: Ref_L
    integer :: Nc_fundamental
    logical :: halfRange
    integer :: mpi_xdim, mpi_ydim, mpi_zdim, block_repx, block_repy, block_repz
    integer :: maxStep, chkConvergeStep, saveStep, saveFormat
    double precision :: eps
    logical :: saveLast
    double precision :: Kn, pressDrop, accom
    character(len=8) :: precision
    integer :: device
    logical :: exact_quadrature, exact_order
    double precision :: refineSwitch
    namelist /physicalNml/ imageFileName, Nx, Ny, Nz, wallExtOrder, fluidlayer, Ref_L
    namelist /velocityNml/ Nc_fundamental, halfRange
    namelist /mpiNml/ mpi_xdim, mpi_ydim, mpi_zdim, block_repx, block_repy, block_repz
    namelist /solverNml/ maxStep, chkConvergeStep, saveStep, eps, saveLast, saveFormat
    namelist /flowNml/ Kn, pressDrop, accom
    namelist /gpuNml/ precision, device, exact_quadrature, exact_order, refineSwitch

    integer, parameter :: PARAFILE = 10, IMAGEFILE = 20
    type(piks_params), target :: prm
    type(piks_info) :: info
    type(c_ptr) :: ctx
    integer(c_int8_t), allocatable, target :: raw(:)
    double precision, allocatable, target :: rho(:), ux(:), uy(:), uz(:)
    integer(c_int32_t), allocatable, target :: labels(:)
    integer :: ios, rc, iStep, nstep, nbytes, next_stop
    double precision :: error, mass, permeability, out(3)
    integer(8) :: t0, t1, trate

    ! ------------------------------------------------------------ para.in
    block_repx = 1; block_repy = 1; block_repz = 1
    wallExtOrder = 2
    fluidlayer = 0
    Ref_L = 1.0d0
    saveLast = .TRUE.
    saveFormat = 1
    mpi_xdim = 1; mpi_ydim = 1; mpi_zdim = 1
    precision = 'fp64'
    device = 0
    exact_quadrature = .false.
    exact_order = .false.
    refineSwitch = 1.0d-6

    open(unit=PARAFILE, file='para.in', status='old', iostat=ios)
    if (ios /= 0) then
        print*, 'ERROR: could not open namelist file'
        stop
    end if
    read(UNIT=PARAFILE, NML=physicalNml, IOSTAT=ios)
    read(UNIT=PARAFILE, NML=velocityNml, IOSTAT=ios)
    read(UNIT=PARAFILE, NML=mpiNml, IOSTAT=ios)
    read(UNIT=PARAFILE, NML=solverNml, IOSTAT=ios)
    read(UNIT=PARAFILE, NML=flowNml, IOSTAT=ios)
    if (ios /= 0) then
        print*, 'ERROR: could not read example namelist'
        stop
    end if
    rewind(PARAFILE)
    read(UNIT=PARAFILE, NML=gpuNml, IOSTAT=ios)   ! optional
    close(PARAFILE)
    if (block_repx /= 1 .or. block_repy /= 1 .or. block_repz /= 1) then
        print*, 'ERROR: block_rep (weak-scaling replication) is not supported'
        stop
    end if
    call print_params()

    ! ------------------------------------------------------------ image
    nbytes = Nx * Ny * Nz
    allocate(raw(nbytes))
    open(IMAGEFILE, file=imageFileName, status='OLD', form='unformatted', access='STREAM', iostat=ios)
    if (ios /= 0) then
        print*, 'ERROR: could not open image file ', trim(imageFileName)
        stop
    end if
    read(IMAGEFILE, iostat=ios) raw
    close(IMAGEFILE)
    if (ios /= 0) then
        write(*,*) "invalid image data file"
        stop 0
    end if

    ! ------------------------------------------------------------ set-up
    prm%nx = Nx; prm%ny = Ny; prm%nz = Nz
    prm%wall_ext_order = wallExtOrder
    prm%fluid_layer = fluidlayer
    prm%ref_l = Ref_L
    prm%nc_fundamental = Nc_fundamental
    prm%half_range = merge(1, 0, halfRange)
    prm%chk_converge_step = chkConvergeStep
    prm%kn = Kn
    prm%press_drop = pressDrop
    prm%accom = accom
    prm%precision = merge(PIKS_PREC_FP32_REFINED, PIKS_PREC_FP64, trim(precision) == 'fp32')
    prm%device = device
    prm%exact_quadrature = merge(1, 0, exact_quadrature)
    prm%exact_order = merge(1, 0, exact_order)
    prm%sweep = 0
    prm%tile = 0
    prm%reserved = 0
    rc = piks_create(prm, c_loc(raw), int(nbytes, c_int64_t), ctx)
    if (rc /= 0) then
        write(*,*) trim(piks_error_message())
        stop 0
    end if
    deallocate(raw)
    rc = piks_get_info(ctx, info)
    print*, "nWall = ", info%nwall
    allocate(rho(info%ntotal), ux(info%ntotal), uy(info%ntotal), uz(info%ntotal), labels(info%ntotal))
    rc = piks_get_labels(ctx, c_loc(labels))

    ! ------------------------------------------------------------ main loop
    ! src/main.F90:95-121; the device runs the steps between two events
    ! (check, save, end) in one call, which changes nothing in the iterates.
    error = 1.d0
    permeability = 0.d0
    mass = 1.d0
    call system_clock(t0, trate)
    iStep = 0
    do while (iStep < maxStep)
        next_stop = min(maxStep, (iStep / chkConvergeStep + 1) * chkConvergeStep, &
                        (iStep / saveStep + 1) * saveStep)
        if (error <= eps .and. iStep < chkConvergeStep) next_stop = iStep + 1
        nstep = next_stop - iStep
        rc = piks_iterate(ctx, int(nstep, c_int32_t))
        call check_rc(rc)
        iStep = next_stop
        if (mod(iStep, chkConvergeStep) == 0) call chk_converge()
        if (mod(iStep, saveStep) == 0) call save_fields()
        if (error <= eps) then
            if (iStep < chkConvergeStep) call chk_converge()
            exit
        endif
    enddo
    if (iStep >= maxStep .and. error > eps) iStep = maxStep + 1
    call system_clock(t1)
    write(*,'(A,ES11.3, A, I6, A, ES15.6)') "Walltime= ", dble(t1 - t0) / dble(trate), &
        ", Steps= ", iStep, ", K= ", permeability

    if (saveLast) call save_fields()
    rc = piks_destroy(ctx)

contains

    subroutine check_rc(code)
        integer, intent(in) :: code
        if (code /= 0) then
            write(*,*) 'piks error ', code, ': ', trim(piks_error_message())
            stop 1
        end if
    end subroutine check_rc

    ! src/output.F90:151-159
    subroutine chk_converge()
        rc = piks_chk_converge(ctx, out)
        call check_rc(rc)
        mass = out(1)
        permeability = out(2)
        error = out(3)
        write(*,"( 1I10, 3ES15.6)") iStep, mass, permeability, error
        open(22, file='Results.dat', position="append")
        write(22,'(4ES15.6, 1I15)') Kn, mass, permeability, error, iStep
        close(22)
    end subroutine chk_converge

    subroutine save_fields()
        rc = piks_get_macro(ctx, c_loc(rho), c_loc(ux), c_loc(uy), c_loc(uz))
        call check_rc(rc)
        select case (saveFormat)
            case (1)
                call save_vti()
            case (2)
                call save_tecplot()
            case (3)
                call save_vtk()
        end select
    end subroutine save_fields

    integer function cell(i, j, k)
        integer, intent(in) :: i, j, k
        cell = (i + 2) + (j + 1) * info%nxt + (k + 1) * info%nxt * info%nyt
    end function cell

    ! src/output.F90:166-189
    subroutine save_tecplot()
        integer :: i, j, k, l
        open(20, file='Field_000.dat', status="REPLACE")
        write(20,*) ' TITLE=" Field"'
        write(20,*) ' VARIABLES=x,y,z,flag,Rho,Ux,Uy,Uz'
        write(20,'(A,A,A,I8,A,I8,A,I8,A)') ' ZONE T=', '"proc000"', ', I=', info%nx, ', J=', &
            info%ny, ', K=', info%nz, ' F=POINT'
        do k = 1, info%nz
            do j = 1, info%ny
                do i = 1, info%nx
                    l = cell(i, j, k)
                    if (labels(l) == 0) then
                        write(20,'(4I10,4ES15.6)') i, j, k, 0, rho(l) + 1.d0, ux(l), uy(l), uz(l)
                    else
                        write(20,'(4I10,4ES15.6)') i, j, k, 1, 0.d0, 0.d0, 0.d0, 0.d0
                    endif
                enddo
            enddo
        enddo
        close(20)
    end subroutine save_tecplot

    ! src/output.F90:192-248
    subroutine save_vtk()
        integer :: i, j, k, l
        open(unit=12, file='Field_000.vtk', status="REPLACE", position="APPEND")
        write(12,'(A)') "# vtk DataFile Version 2.0"
        write(12,'(A)') "DVM MPI"
        write(12,'(A)') "ASCII"
        write(12,*)
        write(12,'(A)') "DATASET STRUCTURED_POINTS"
        write(12,*) "DIMENSIONS", info%nx, info%ny, info%nz
        write(12,*) "ORIGIN", 1, 1, 1
        write(12,*) "SPACING", 1, 1, 1
        write(12,*)
        write(12,*) "POINT_DATA", info%nx * info%ny * info%nz
        write(12,*)
        write(12,'(A)') "SCALARS Rho double"
        write(12,'(A)') "LOOKUP_TABLE default"
        do k = 1, info%nz
            do j = 1, info%ny
                do i = 1, info%nx
                    l = cell(i, j, k)
                    if (labels(l) == 0) then
                        write(12,'(ES15.6)') rho(l) + 1.d0
                    else
                        write(12,'(ES15.6)') 0.d0
                    endif
                enddo
            enddo
        enddo
        write(12,*)
        write(12,'(A)') "VECTORS Velocity double"
        do k = 1, info%nz
            do j = 1, info%ny
                do i = 1, info%nx
                    l = cell(i, j, k)
                    if (labels(l) == 0) then
                        write(12,'(3ES15.6)') ux(l), uy(l), uz(l)
                    else
                        write(12,'(3ES15.6)') 0.d0, 0.d0, 0.d0
                    endif
                enddo
            enddo
        enddo
        close(unit=12)
    end subroutine save_vtk

    ! src/output.F90:250-360 (one piece: the whole domain)
    subroutine save_vti()
        integer :: i, j, k, l
        open(unit=13, file='Field_000.vti', status="REPLACE", position="APPEND")
        write(13,'(A)') '<?xml version="1.0"?>'
        write(13,'(A)') '<VTKFile type="ImageData">'
        write(13,'(A, 6I4, A)') '<ImageData WholeExtent="', 0, info%nx, 0, info%ny, 0, info%nz, &
            ' " Origin="0 0 0" Spacing="1 1 1">'
        write(13,'(A, 6I4, A)') '<Piece Extent="', 0, info%nx, 0, info%ny, 0, info%nz, '">'
        write(13,'(A)') '<CellData Scalars="flag Rho" Vectors ="U">'
        write(13,'(A)') '<DataArray type="Int32" Name="flag" format="ascii">'
        do k = 1, info%nz
            do j = 1, info%ny
                do i = 1, info%nx
                    write(13,'(I2)') labels(cell(i, j, k))
                enddo
            enddo
        enddo
        write(13,'(A)') '</DataArray>'
        write(13,'(A)') '<DataArray type="Float32" Name="Rho" format="ascii">'
        do k = 1, info%nz
            do j = 1, info%ny
                do i = 1, info%nx
                    l = cell(i, j, k)
                    if (labels(l) == 0) then
                        write(13,'(ES15.6)') rho(l) + 1.d0
                    else
                        write(13,'(ES15.6)') 0.d0
                    endif
                enddo
            enddo
        enddo
        write(13,'(A)') '</DataArray>'
        write(13,'(A)') '<DataArray type="Float32" Name="U" format="ascii"  NumberOfComponents="3">'
        do k = 1, info%nz
            do j = 1, info%ny
                do i = 1, info%nx
                    l = cell(i, j, k)
                    if (labels(l) == 0) then
                        write(13,'(3ES15.6)') ux(l), uy(l), uz(l)
                    else
                        write(13,'(3ES15.6)') 0.d0, 0.d0, 0.d0
                    endif
                enddo
            enddo
        enddo
        write(13,'(A)') '</DataArray>'
        write(13,'(A)') '</CellData>'
        write(13,'(A)') '</Piece>'
        write(13,'(A)') '</ImageData>'
        write(13,'(A)') '</VTKFile>'
        close(unit=13)

        open(unit=14, file='Field.pvti', status="REPLACE", position="APPEND")
        write(14,'(A)') '<?xml version="1.0"?>'
        write(14,'(A)') '<VTKFile type="PImageData">'
        write(14,'(A, 6I4, A)') '<PImageData WholeExtent="', 0, info%nx, 0, info%ny, 0, info%nz, &
            ' " Origin="0 0 0" Spacing="1 1 1">'
        write(14,'(A)') '<PCellData Scalars="flag Rho" Vectors ="U">'
        write(14,'(A)') '<DataArray type="Int32" Name="flag" format="ascii"/>'
        write(14,'(A)') '<DataArray type="Float32" Name="Rho" format="ascii"/>'
        write(14,'(A)') '<DataArray type="Float32" Name="U" format="ascii"  NumberOfComponents="3"/>'
        write(14,'(A)') '</PCellData>'
        write(14,'(A, 6I4, A)') '<Piece Extent="', 0, info%nx, 0, info%ny, 0, info%nz, &
            '" Source="Field_000.vti"/>'
        write(14,'(A)') '</PImageData>'
        write(14,'(A)') '</VTKFile>'
        close(unit=14)
    end subroutine save_vti

    ! src/parameters.F90:139-168
    subroutine print_params()
        print*, "========== Parameters ================"
        print*, "imageFileName = ", imageFileName
        print*, "Nx = ", Nx + 2 * fluidlayer
        print*, "Ny = ", Ny
        print*, "Nz = ", Nz
        print*, "wallExtOrder = ", wallExtOrder
        print*, "fluidlayer = ", fluidlayer
        print*, "Ref_L = ", Ref_L
        print*, "Nc_fundamental = ", Nc_fundamental
        print*, "halfRange = ", halfRange
        print*, "mpi_xdim = ", mpi_xdim
        print*, "mpi_ydim = ", mpi_ydim
        print*, "mpi_zdim = ", mpi_zdim
        print*, "block_repx = ", block_repx
        print*, "block_repy = ", block_repy
        print*, "block_repz = ", block_repz
        print*, "maxStep = ", maxStep
        print*, "chkConvergeStep = ", chkConvergeStep
        print*, "saveStep = ", saveStep
        print*, "eps = ", eps
        print*, "saveLast = ", saveLast
        print*, "saveFormat = ", saveFormat
        print*, "Kn = ", Kn
        print*, "pressDrop = ", pressDrop
        print*, "accom = ", accom
        print*, "precision = ", trim(precision)
        print*, "======================================"
    end subroutine print_params

end program piks3d_gpu
