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
!                bounded sample for the CPU baseline.
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

   open(77, file=trim(dfile), access='stream', form='unformatted', status='old', iostat=ios)
   if (ios /= 0) stop 'cannot open density file'
   read(77) quick_qm_struct%dense
   close(77)

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
   else if (trim(mode) == 'grad') then
      quick_qm_struct%gradient = 0.0d0
      call cpu_time(t0)
      ! density screening and four-centre loop exactly as gradient.f90:266-320
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
      call cpu_time(t1)
      call dump_d(trim(odir)//'/grad4c.bin', quick_qm_struct%gradient, 3*natom)
      call dump_d(trim(odir)//'/time_grad.bin', (/ t1-t0 /), 1)
   endif

contains

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
