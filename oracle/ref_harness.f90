! ref_harness.f90 -- test driver (ours) around the reference's own FFT-free
! subroutines, compiled from the reference's sources by oracle/build_ref.sh in FP64
! (-fdefault-real-8 -fdefault-double-8).  TEST INFRASTRUCTURE ONLY.
!
! mode 1: feeds the 3/2-grid gradients, expanded level-1/nnz velocities and
!         SGS TKE of every level into the compbuf rings exactly as comp1 does
!         (comp1.f90:67-89), calls the reference compsgs1 (which calls complen,
!         sufsim, sufsim_upper) for iz = 1..nnz and writes the stresses and
!         eddy diffusivities it leaves in the rings.
! mode 2: solves the tridiagonal systems of the input with the reference tridv.
! mode 3: loads u, v (levels 1..nnz+1), w (faces 0..nnz), e (0..nnz+1) and p
!         (1..nnz+1), calls the reference compmn(1, nnz) and writes meanbuf.
!
! Stream files, native endianness: input on unit 10 (argv 1), output unit 11
! (argv 2).
program ref_harness
  use storage
  use compbuf
  use parameters
  implicit none
  integer :: mode, np(4), nptot(4), nm(4), nmtot(4)
  character(len=512) :: fin, fout
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(10, file=trim(fin), access="stream", form="unformatted", status="old")
  open(11, file=trim(fout), access="stream", form="unformatted", status="replace")
  read(10) mode
  if (mode == 1) then
     call run_sgs()
  else if (mode == 2) then
     call run_tridv()
  else if (mode == 3) then
     call run_means()
  end if
  close(10)
  close(11)

contains

  subroutine run_sgs()
    integer :: im, isf, iz, j, jp1, k, km1, l, lm1, lp1, nm1, n, np1, np2, kb
    real :: alkv
    real, allocatable :: uxa(:,:,:), uya(:,:,:), vxa(:,:,:), vya(:,:,:), wza(:,:,:)
    real, allocatable :: uza(:,:,:), vza(:,:,:), wxa(:,:,:), wya(:,:,:), eba(:,:,:)
    real, allocatable :: ub1(:,:), ubn(:,:), vb1(:,:), vbn(:,:)
    read(10) nnx, nny, nnz, im, isf
    read(10) xl, yl, zl, z0, dt, alkv
    imodel = im
    isurface = isf
    isbgr = .true.
    spectav = .true.
    ! derived constants as the driver computes them (les.F90:99-121, :216)
    nnxb = 3*nnx/2
    nnyb = 3*nny/2
    xymb = 1.0/float(nnxb)/float(nnyb)
    emin = 1.e-6
    ncx = nnx/2+1
    ncy = nny/2+1
    nxy = nnx*nny
    xym = 1./float(nnx*nny)
    dx = xl/float(nnx)
    dy = yl/float(nny)
    dz = zl/float(nnz)
    z1 = dz/2.
    zody = log(z1/z0)
    dzm = 1./dz
    dzm2 = dzm**2
    todzm = 0.5/dz
    dsl = (dx*dy*dz)**(1./3.)
    ceps = 0.93/dsl
    xkmax = dz*dz/dt/5.
    ugal = 0.0
    vgal = 0.0
    call setcombuf(np, nptot, nm, nmtot)
    call allocate_arrays(1, nnz, np, nm)
    alk = alkv
    allocate(uxa(nnxb,nnyb,0:nnz+1), uya(nnxb,nnyb,0:nnz+1), vxa(nnxb,nnyb,0:nnz+1))
    allocate(vya(nnxb,nnyb,0:nnz+1), wza(nnxb,nnyb,0:nnz+1))
    allocate(uza(nnxb,nnyb,0:nnz), vza(nnxb,nnyb,0:nnz), wxa(nnxb,nnyb,0:nnz))
    allocate(wya(nnxb,nnyb,0:nnz), eba(nnxb,nnyb,0:nnz))
    allocate(ub1(nnxb,nnyb), ubn(nnxb,nnyb), vb1(nnxb,nnyb), vbn(nnxb,nnyb))
    read(10) uxa, uya, vxa, vya, wza
    read(10) uza, vza, wxa, wya, eba
    read(10) ub1, ubn, vb1, vbn
    read(10) ru(:,:,1:nnz), rv(:,:,1:nnz)
    kb = 0
    do iz = 1, nnz
       ! ring pointers of comp1.f90:70-89
       j = mod(iz,2)+1
       jp1 = mod(iz+1,2)+1
       k = mod(iz,2)+1
       km1 = mod(iz+1,2)+1
       lm1 = mod(iz+2,3)+1
       l = mod(iz,3)+1
       lp1 = mod(iz+1,3)+1
       nm1 = mod(iz+3,4)+1
       n = mod(iz,4)+1
       np1 = mod(iz+1,4)+1
       np2 = mod(iz+2,4)+1
       ux(:,:,nm1) = uxa(:,:,iz-1); ux(:,:,n) = uxa(:,:,iz)
       ux(:,:,np1) = uxa(:,:,iz+1); ux(:,:,np2) = uxa(:,:,min(iz+2,nnz+1))
       uy(:,:,nm1) = uya(:,:,iz-1); uy(:,:,n) = uya(:,:,iz)
       uy(:,:,np1) = uya(:,:,iz+1); uy(:,:,np2) = uya(:,:,min(iz+2,nnz+1))
       vx(:,:,nm1) = vxa(:,:,iz-1); vx(:,:,n) = vxa(:,:,iz)
       vx(:,:,np1) = vxa(:,:,iz+1); vx(:,:,np2) = vxa(:,:,min(iz+2,nnz+1))
       vy(:,:,nm1) = vya(:,:,iz-1); vy(:,:,n) = vya(:,:,iz)
       vy(:,:,np1) = vya(:,:,iz+1); vy(:,:,np2) = vya(:,:,min(iz+2,nnz+1))
       wz(:,:,nm1) = wza(:,:,iz-1); wz(:,:,n) = wza(:,:,iz)
       wz(:,:,np1) = wza(:,:,iz+1); wz(:,:,np2) = wza(:,:,min(iz+2,nnz+1))
       uz(:,:,lm1) = uza(:,:,iz-1); uz(:,:,l) = uza(:,:,iz); uz(:,:,lp1) = uza(:,:,min(iz+1,nnz))
       vz(:,:,lm1) = vza(:,:,iz-1); vz(:,:,l) = vza(:,:,iz); vz(:,:,lp1) = vza(:,:,min(iz+1,nnz))
       wx(:,:,lm1) = wxa(:,:,iz-1); wx(:,:,l) = wxa(:,:,iz); wx(:,:,lp1) = wxa(:,:,min(iz+1,nnz))
       wy(:,:,lm1) = wya(:,:,iz-1); wy(:,:,l) = wya(:,:,iz); wy(:,:,lp1) = wya(:,:,min(iz+1,nnz))
       ebuf(:,:,-1) = eba(:,:,iz-1); ebuf(:,:,0) = eba(:,:,iz); ebuf(:,:,1) = eba(:,:,min(iz+1,nnz))
       ubuf(:,:,0) = 0.0
       vbuf(:,:,0) = 0.0
       if (iz == 1) then
          ubuf(:,:,0) = ub1
          vbuf(:,:,0) = vb1
       else if (iz == nnz) then
          ubuf(:,:,0) = ubn
          vbuf(:,:,0) = vbn
       end if
       call compsgs1(kb,iz,j,jp1,k,km1,l,lm1,lp1,nm1,n,np1,np2,1)
       if (iz == 1) write(11) tauuw(:,:,km1), tauvw(:,:,km1), xkk(:,:,lm1)
       write(11) tauuu, tauvv, tauuv, tauww(:,:,j), tauuw(:,:,k), tauvw(:,:,k), xkk(:,:,l)
       if (iz == 1) write(11) utau, utau_s, dudt, dvdt
    end do
    write(11) utau
  end subroutine run_sgs

  subroutine run_tridv()
    integer :: ny, nz
    real, allocatable :: bdar(:,:,:)
    read(10) ny, nz
    allocate(bdar(ny,nz,4))
    read(10) bdar
    call tridv(bdar, ny, nz)
    write(11) bdar(:,:,4)
  end subroutine run_tridv

  subroutine run_means()
    integer :: np(4), nptot(4), nm(4), nmtot(4)
    read(10) nnx, nny, nnz
    read(10) xl, yl, zl
    isbgr = .true.
    spectav = .true.
    nnxb = 3*nnx/2
    nnyb = 3*nny/2
    xymb = 1.0/float(nnxb)/float(nnyb)
    ncx = nnx/2+1
    ncy = nny/2+1
    nxy = nnx*nny
    xym = 1./float(nnx*nny)
    call setcombuf(np, nptot, nm, nmtot)
    call allocate_arrays(1, nnz, np, nm)
    u = 0.0
    v = 0.0
    w = 0.0
    e = 0.0
    p = 0.0
    read(10) u(:,:,1:nnz+1), v(:,:,1:nnz+1), w(:,:,0:nnz), e(:,:,0:nnz+1), p(1:nnx,:,1:nnz+1)
    call compmn(1, nnz)
    write(11) meanbuf(:,1:nnz)
  end subroutine run_means

end program ref_harness
