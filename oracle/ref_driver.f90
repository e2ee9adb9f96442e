! Golden-vector driver for the reference CPU path (test infrastructure only).
!
! Linked against the reference sources compiled by oracle/ref_build.mk.  It runs
! the reference's own set-up sequence (main.f90:53-164: initialize1,
! set_quick_files, read_Job_and_Atom, getMol, g2eshell, schwarzoff), then, for a
! density read from a file:
!   mode "fock": hfoperator with a zero one-electron operator
!                (hfoperator.f90:24-138), i.e. the two-electron part G(D);
!   mode "grad": the four-centre gradient loop of hfgrad (gradient.f90:261-320);
!   mode "fock_sample <stride>": the CPU Fock build's outer loop (hfoperator.f90:
!                109-111) restricted to every stride-th shell II, timed -- a
!                bounded sample for the CPU baseline;
!   mode "scf":  a restricted SCF to a converged density: the reference's own
!                overlap/orthogonaliser (fullX), one-electron operator (get1e),
!                Fock build (hfoperator) and diagonaliser (DIAG), iterated by
!                Pulay DIIS on the error FDS - SDF (as electdiis, scf.f90:170-537,
!                whose per-cycle checkpoint writer LLVM flang's runtime rejects).
!                From the reference's initial guess (getMol) it runs until the
!                density's RMS change is below 1e-10, then dumps D, H_core,
!                G(D) (hfoperator with a zero one-electron part), the nuclear
!                repulsion (getEnergy.f90:31-52) and E = sum D H + 1/2 sum D G
!                + E_nn (hfoperator.f90:791-803, getEnergy.f90:113), and the
!                occupied orbitals and their energies;
!   mode "scfgrad": "scf", then at the converged density the full gradient of
!                hfgrad (gradient.f90:5-327) and its four-centre term alone;
!   mode "xc":   (job card B3LYP) the reference's own SG-1 grid -- every point
!                dftoperator visits, enumerated as dft.f90:906-936 does with
!                gridformnew, SSW and the radial tables of gridformSG1, with its
!                weight when that is >= DMCutoff -- and the significance radii
!                of get_sigrad; then dftoperator for the given density
!                (dft.f90:618-1553): the final operator (1e + 2e with exchange
!                weight 0.2 + XC) and Osavedft, the operator before the XC
!                terms, whose difference is V_xc (lower triangle); alpha and
!                beta electron counts.  dftoperator prints E_xc ("Eelxc =").
! Everything is dumped as raw little-endian binary so the Python fixture writer
! can read it without running any reference code itself.
!
! usage: qk_ref_driver <job.in> <mode> <density.bin> <outdir> [stride]
program qk_ref_driver
   use allmod
   implicit none
   integer :: ierr, i, j, nb, ios, stride, nsamp
   character(len=32) :: sarg
   character(len=256) :: mode, dfile, odir
   double precision, allocatable :: oneElecO(:,:)
   double precision :: Testtmp, testCutoff, cutoffTest, DNtemp, t0, t1
   double precision, allocatable :: hold(:,:), hold2(:,:), Fh(:,:,:), Eh(:,:,:), Bm(:,:), cf(:), V2(:,:), Dold(:,:)
   double precision :: Enn, Eel2, Etot, prms, pchg, dist
   integer :: it, nd, k, nocc, ierr2, nsave
   integer :: Iatm, Irad, Iang, iiangt, npts
   double precision :: rad, rad3, gx, gy, gz, wpt
   double precision, external :: ssw
   integer II,JJ,KK,LL,NBI1,NBI2,NBJ1,NBJ2,NBK1,NBK2,NBL1,NBL2
   common /hrrstore/II,JJ,KK,LL,NBI1,NBI2,NBJ1,NBJ2,NBK1,NBK2,NBL1,NBL2

   call getarg(2, mode)
   call getarg(3, dfile)
   call getarg(4, odir)

   call initialize1(ierr)
   call set_quick_files(ierr)
   call quick_open(iOutFile,outFileName,'U','F','R',.false.)
   call read_Job_and_Atom()
   call alloc(quick_molspec)
   if (quick_method%SAD) call getMolSad()
   call getMol()
   call g2eshell
   call schwarzoff

   nb = nbasis
   call dump_i(trim(odir)//'/sizes.bin', (/ natom, nbasis, nshell, jshell, jbasis, MAXPRIM /), 6)
   call dump_d(trim(odir)//'/xyz.bin', xyz, 3*natom)
   call dump_i(trim(odir)//'/kstart.bin', quick_basis%kstart, nshell)
   call dump_i(trim(odir)//'/kprim.bin', quick_basis%kprim, nshell)
   call dump_i(trim(odir)//'/katom.bin', quick_basis%katom, nshell)
   call dump_i(trim(odir)//'/ktype.bin', quick_basis%ktype, nshell)
   call dump_i(trim(odir)//'/Qstart.bin', quick_basis%Qstart, nshell)
   call dump_i(trim(odir)//'/Qfinal.bin', quick_basis%Qfinal, nshell)
   call dump_i(trim(odir)//'/Qsbasis.bin', quick_basis%Qsbasis, nshell*4)
   call dump_i(trim(odir)//'/Qfbasis.bin', quick_basis%Qfbasis, nshell*4)
   call dump_i(trim(odir)//'/ksumtype.bin', quick_basis%ksumtype, nshell+1)
   call dump_i(trim(odir)//'/KLMN.bin', quick_basis%KLMN, 3*nb)
   call dump_d(trim(odir)//'/cons.bin', quick_basis%cons, nb)
   call dump_d(trim(odir)//'/gccoeff.bin', quick_basis%gccoeff, MAXPRIM*nb)
   call dump_d(trim(odir)//'/gcexpo.bin', quick_basis%gcexpo, MAXPRIM*nb)
   call dump_d(trim(odir)//'/Ycutoff.bin', Ycutoff, nshell*nshell)
   call dump_d(trim(odir)//'/cutprim.bin', cutprim, jbasis*jbasis)

   if (trim(mode) == 'setup') stop

   if (trim(mode) /= 'scf' .and. trim(mode) /= 'scfgrad') then
      open(77, file=trim(dfile), access='stream', form='unformatted', status='old', iostat=ios)
      if (ios /= 0) stop 'cannot open density file'
      read(77) quick_qm_struct%dense
      close(77)
   endif

   if (trim(mode) == 'fock') then
      allocate(oneElecO(nb,nb))
      oneElecO = 0.0d0
      quick_method%printEnergy = .false.
      call cpu_time(t0)
      call hfoperator(oneElecO, .false.)
      call cpu_time(t1)
      call dump_d(trim(odir)//'/G.bin', quick_qm_struct%o, nb*nb)
      call dump_d(trim(odir)//'/time_fock.bin', (/ t1-t0 /), 1)
   else if (trim(mode) == 'fock_sample') then
      call getarg(5, sarg)
      read(sarg, *) stride
      quick_qm_struct%o = 0.0d0
      call densityCutoff()
      nsamp = 0
      call cpu_time(t0)
      do II = max(1, stride / 2), jshell, stride
         call get2e(II)
         nsamp = nsamp + 1
      enddo
      call cpu_time(t1)
      call dump_d(trim(odir)//'/time_sample.bin', (/ t1-t0, dble(nsamp), dble(jshell) /), 3)
   else if (trim(mode) == 'scf' .or. trim(mode) == 'scfgrad') then
      allocate(oneElecO(nb,nb), hold(nb,nb), hold2(nb,nb), Fh(nb,nb,8), Eh(nb,nb,8), Bm(9,9), cf(9), &
               V2(3,nb), Dold(nb,nb))
      call fullX
      Enn = 0.0d0
      do i = 1, natom
         do j = i+1, natom
            dist = sqrt(sum((xyz(1:3,i) - xyz(1:3,j))**2))
            Enn = Enn + quick_molspec%chg(i)*quick_molspec%chg(j)/dist
         enddo
      enddo
      call get1e(oneElecO)
      quick_method%printEnergy = .true.
      nocc = quick_molspec%nelec/2
      nsave = 0
      call cpu_time(t0)
      do it = 1, 200
         call hfoperator(oneElecO, .false.)
         ! DIIS: error e = F D S - S D F, up to eight previous operators
         call DGEMM('n','n',nb,nb,nb,1.0d0,quick_qm_struct%dense,nb,quick_qm_struct%s,nb,0.0d0,hold,nb)
         call DGEMM('n','n',nb,nb,nb,1.0d0,quick_qm_struct%o,nb,hold,nb,0.0d0,hold2,nb)
         k = mod(nsave, 8) + 1
         nsave = nsave + 1
         nd = min(nsave, 8)
         Fh(:,:,k) = quick_qm_struct%o
         Eh(:,:,k) = hold2 - transpose(hold2)
         do i = 1, nd
            do j = 1, nd
               Bm(i,j) = sum(Eh(:,:,i)*Eh(:,:,j))
            enddo
            Bm(i,nd+1) = -1.0d0
            Bm(nd+1,i) = -1.0d0
         enddo
         Bm(nd+1,nd+1) = 0.0d0
         cf = 0.0d0
         cf(nd+1) = -1.0d0
         call gauss_solve(Bm, cf, nd+1, ierr2)
         if (ierr2 == 0) then
            quick_qm_struct%o = 0.0d0
            do i = 1, nd
               quick_qm_struct%o = quick_qm_struct%o + cf(i)*Fh(:,:,i)
            enddo
         endif
         call DGEMM('n','n',nb,nb,nb,1.0d0,quick_qm_struct%o,nb,quick_qm_struct%x,nb,0.0d0,hold,nb)
         call DGEMM('n','n',nb,nb,nb,1.0d0,quick_qm_struct%x,nb,hold,nb,0.0d0,quick_qm_struct%o,nb)
         call DIAG(nb,quick_qm_struct%o,nb,quick_method%DMCutoff,V2,quick_qm_struct%E, &
                   quick_qm_struct%idegen,quick_qm_struct%vec,ierr2)
         call DGEMM('n','n',nb,nb,nb,1.0d0,quick_qm_struct%x,nb,quick_qm_struct%vec,nb,0.0d0, &
                    quick_qm_struct%co,nb)
         Dold = quick_qm_struct%dense
         call DGEMM('n','t',nb,nb,nocc,2.0d0,quick_qm_struct%co,nb,quick_qm_struct%co,nb,0.0d0, &
                    quick_qm_struct%dense,nb)
         pchg = maxval(abs(quick_qm_struct%dense - Dold))
         prms = sqrt(sum((quick_qm_struct%dense - Dold)**2)/dble(nb*nb))
         write(*,'(a,i4,a,f22.12,a,es10.3,a,es10.3)') ' scf', it, ' E', quick_qm_struct%Eel + Enn, &
               ' rms', prms, ' max', pchg
         if (prms < 1.0d-10 .and. pchg < 1.0d-8) exit
      enddo
      call cpu_time(t1)
      ! the energy functional and G at the final density
      hold = oneElecO
      oneElecO = 0.0d0
      quick_method%printEnergy = .false.
      call hfoperator(oneElecO, .false.)
      oneElecO = hold
      Eel2 = sum(quick_qm_struct%dense*oneElecO) + 0.5d0*sum(quick_qm_struct%dense*quick_qm_struct%o)
      Etot = Eel2 + Enn
      call dump_d(trim(odir)//'/D.bin', quick_qm_struct%dense, nb*nb)
      call dump_d(trim(odir)//'/H.bin', oneElecO, nb*nb)
      call dump_d(trim(odir)//'/S.bin', quick_qm_struct%s, nb*nb)
      call dump_d(trim(odir)//'/G.bin', quick_qm_struct%o, nb*nb)
      call dump_d(trim(odir)//'/energy.bin', (/ Etot, Eel2, Enn, dble(it), prms, t1-t0 /), 6)
      call dump_d(trim(odir)//'/C.bin', quick_qm_struct%co, nb*nb)
      call dump_d(trim(odir)//'/eps.bin', quick_qm_struct%E, nb)
      if (trim(mode) == 'scfgrad') then
         ! the full analytic gradient (gradient.f90:5-327: nuclear repulsion,
         ! overlap x W, kinetic, nuclear attraction and four-centre terms), then
         ! the four-centre term alone
         call cpu_time(t0)
         call hfgrad
         call cpu_time(t1)
         call dump_d(trim(odir)//'/gradtotal.bin', quick_qm_struct%gradient, 3*natom)
         call four_centre()
         call dump_d(trim(odir)//'/grad4c.bin', quick_qm_struct%gradient, 3*natom)
         call dump_d(trim(odir)//'/time_grad.bin', (/ t1-t0 /), 1)
      endif
   else if (trim(mode) == 'xc') then
      open(79, file=trim(odir)//'/grid.bin', access='stream', form='unformatted', status='replace')
      npts = 0
      do Iatm = 1, natom
         do Irad = 1, 50
            call gridformnew(Iatm, RGRID(Irad), iiangt)
            rad = radii(quick_molspec%iattype(Iatm))
            rad3 = rad*rad*rad
            do Iang = 1, iiangt
               gx = xyz(1,Iatm) + rad*RGRID(Irad)*XANG(Iang)
               gy = xyz(2,Iatm) + rad*RGRID(Irad)*YANG(Iang)
               gz = xyz(3,Iatm) + rad*RGRID(Irad)*ZANG(Iang)
               wpt = ssw(gx, gy, gz, Iatm)*WTANG(Iang)*RWT(Irad)*rad3
               if (wpt >= quick_method%DMCutoff) then
                  write(79) gx, gy, gz, wpt, dble(Iatm)
                  npts = npts + 1
               endif
            enddo
         enddo
      enddo
      close(79)
      call dump_i(trim(odir)//'/npts.bin', (/ npts /), 1)
      call dump_d(trim(odir)//'/sigrad2.bin', sigrad2, nb)
      quick_method%printEnergy = .true.
      call cpu_time(t0)
      call dftoperator
      call cpu_time(t1)
      call dump_d(trim(odir)//'/O.bin', quick_qm_struct%o, nb*nb)
      call dump_d(trim(odir)//'/Osave.bin', quick_qm_struct%Osavedft, nb*nb)
      call dump_d(trim(odir)//'/xc_scalars.bin', (/ quick_qm_struct%aelec, quick_qm_struct%belec, &
                  quick_qm_struct%Eel, t1-t0 /), 4)
   else if (trim(mode) == 'grad') then
      call cpu_time(t0)
      call four_centre()
      call cpu_time(t1)
      call dump_d(trim(odir)//'/grad4c.bin', quick_qm_struct%gradient, 3*natom)
      call dump_d(trim(odir)//'/time_grad.bin', (/ t1-t0 /), 1)
   endif

contains

   ! density screening and four-centre loop exactly as gradient.f90:266-320
   subroutine four_centre()
      quick_qm_struct%gradient = 0.0d0
      do II=1,jshell
         do JJ=II,jshell
            DNtemp=0.0d0
            call DNscreen(II,JJ,DNtemp)
            Cutmatrix(II,JJ)=DNtemp
            Cutmatrix(JJ,II)=DNtemp
         enddo
      enddo
      do II=1,jshell
         do JJ=II,jshell
            Testtmp=Ycutoff(II,JJ)
            do KK=II,jshell
               do LL=KK,jshell
                  if(quick_basis%katom(II).eq.quick_basis%katom(JJ).and.quick_basis%katom(II).eq. &
                     quick_basis%katom(KK).and.quick_basis%katom(II).eq.quick_basis%katom(LL))then
                     continue
                  else
                     testCutoff = Testtmp*Ycutoff(KK,LL)
                     if(testCutoff.gt.quick_method%gradCutoff)then
                        DNmax=max(4.0d0*cutmatrix(II,JJ),4.0d0*cutmatrix(KK,LL), &
                             cutmatrix(II,LL),cutmatrix(II,KK),cutmatrix(JJ,KK),cutmatrix(JJ,LL))
                        cutoffTest=testCutoff*DNmax
                        if(cutoffTest.gt.quick_method%gradCutoff)then
                           call shellopt
                        endif
                     endif
                  endif
               enddo
            enddo
         enddo
      enddo
   end subroutine four_centre

   ! Gaussian elimination with partial pivoting: solves B x = rhs in place (rhs <- x)
   subroutine gauss_solve(B, rhs, m, info)
      integer :: m, info
      double precision :: B(9,9), rhs(9), A(9,9), t, piv
      integer :: r, c2, p2
      A = B
      info = 0
      do c2 = 1, m
         p2 = c2
         do r = c2+1, m
            if (abs(A(r,c2)) > abs(A(p2,c2))) p2 = r
         enddo
         if (abs(A(p2,c2)) < 1.0d-300) then
            info = 1
            return
         endif
         if (p2 /= c2) then
            do r = 1, m
               t = A(c2,r); A(c2,r) = A(p2,r); A(p2,r) = t
            enddo
            t = rhs(c2); rhs(c2) = rhs(p2); rhs(p2) = t
         endif
         do r = c2+1, m
            piv = A(r,c2)/A(c2,c2)
            A(r,c2:m) = A(r,c2:m) - piv*A(c2,c2:m)
            rhs(r) = rhs(r) - piv*rhs(c2)
         enddo
      enddo
      do r = m, 1, -1
         rhs(r) = (rhs(r) - sum(A(r,r+1:m)*rhs(r+1:m)))/A(r,r)
      enddo
   end subroutine gauss_solve

   subroutine dump_d(fname, a, n)
      character(len=*) :: fname
      integer :: n
      double precision :: a(n)
      open(78, file=fname, access='stream', form='unformatted', status='replace')
      write(78) a
      close(78)
   end subroutine dump_d

   subroutine dump_i(fname, a, n)
      character(len=*) :: fname
      integer :: n
      integer :: a(n)
      open(78, file=fname, access='stream', form='unformatted', status='replace')
      write(78) a
      close(78)
   end subroutine dump_i

end program qk_ref_driver
