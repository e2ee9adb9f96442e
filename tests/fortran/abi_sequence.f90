! QUICK's Fortran call sequence against libquick_mi355x.so (test program).
!
! Written from scratch to exercise the drop-in boundary the way QUICK's own
! Fortran does: implicit interfaces, every argument by reference, column-major
! arrays with 1-based index contents.  It reads a system, a density, its shell
! block maxima, an operator to accumulate into and a gradient seed from a
! stream file written by the test harness, then calls, in order:
!   main.f90:84-177       gpu_startup, gpu_set_device, gpu_init, gpu_setup,
!                         gpu_upload_xyz, gpu_upload_atom_and_chg,
!                         gpu_upload_basis, gpu_upload_cutoff_matrix
!   hfoperator.f90:77-103 gpu_upload_method, gpu_upload_calculated,
!                         gpu_upload_cutoff (three arguments), gpu_get2e
!   gradient.f90:279-287  gpu_upload_calculated, gpu_upload_cutoff,
!                         gpu_upload_grad, gpu_grad
! and gpu_cleanup, gpu_shutdown; it writes the operator and the gradient.
!
! usage: abi_sequence <input.bin> <output.bin>
program abi_sequence
   implicit none
   integer :: natom, nbasis, nshell, jbasis, maxprim, nelec, method
   integer :: imult, molchg, gpu_id, nprim, jshell, maxcontract
   double precision :: integralCutoff, primLimit, gradCutoff
   double precision, allocatable :: xyz(:,:), chg(:), cons(:), gccoeff(:,:), gcexpo(:,:), &
         Ycutoff(:,:), cutprim(:,:), dense(:,:), cutmatrix(:,:), o(:,:), co(:,:), vec(:,:), &
         grad(:), aexp(:,:), dcoeff(:,:)
   integer, allocatable :: iattype(:), kstart(:), kprim(:), katom(:), ktype(:), Qstart(:), &
         Qfinal(:), Qsbasis(:,:), Qfbasis(:,:), ksumtype(:), KLMN(:,:), ncontract(:), itype(:,:), &
         first_bf(:), last_bf(:), first_shell_bf(:), last_shell_bf(:), ncenter(:), kshell(:), Qnumber(:)
   character(len=512) :: fin, fout

   call getarg(1, fin)
   call getarg(2, fout)
   open(10, file=trim(fin), access='stream', form='unformatted', status='old')
   read(10) natom, nbasis, nshell, jbasis, maxprim, nelec, method
   allocate(xyz(3,natom), chg(natom), iattype(natom), kstart(nshell), kprim(nshell), katom(nshell), &
            ktype(nshell), Qstart(nshell), Qfinal(nshell), Qsbasis(nshell,0:3), Qfbasis(nshell,0:3), &
            ksumtype(nshell+1), KLMN(3,nbasis), cons(nbasis), gccoeff(maxprim,nbasis), &
            gcexpo(maxprim,nbasis), Ycutoff(nshell,nshell), cutprim(jbasis,jbasis), dense(nbasis,nbasis), &
            cutmatrix(nshell,nshell), o(nbasis,nbasis), grad(3*natom))
   read(10) xyz, iattype, chg, kstart, kprim, katom, ktype, Qstart, Qfinal, Qsbasis, Qfbasis, ksumtype, &
            KLMN, cons, gccoeff, gcexpo, Ycutoff, cutprim, dense, cutmatrix, o, grad
   close(10)
   ! arrays of gpu_upload_basis that the library does not read
   allocate(co(nbasis,nbasis), vec(nbasis,nbasis), aexp(maxprim,nbasis), dcoeff(maxprim,nbasis), &
            ncontract(nbasis), itype(3,nbasis), first_bf(natom), last_bf(natom), first_shell_bf(natom), &
            last_shell_bf(natom), ncenter(nbasis), kshell(natom), Qnumber(nshell))
   co = 0.0d0; vec = 0.0d0; aexp = 0.0d0; dcoeff = 0.0d0
   ncontract = 0; itype = 0; first_bf = 0; last_bf = 0; first_shell_bf = 0; last_shell_bf = 0
   ncenter = 0; kshell = 0; Qnumber = 0
   nprim = jbasis; jshell = nshell; maxcontract = maxprim
   imult = 1; molchg = 0; gpu_id = -1
   integralCutoff = 1.0d-11; primLimit = 1.0d-20; gradCutoff = 1.0d-11

   call gpu_startup()
   call gpu_set_device(gpu_id)
   call gpu_init()
   call gpu_setup(natom, nbasis, nelec, imult, molchg, iattype)
   call gpu_upload_xyz(xyz)
   call gpu_upload_atom_and_chg(iattype, chg)
   call gpu_upload_basis(nshell, nprim, jshell, jbasis, maxcontract, ncontract, itype, aexp, dcoeff, &
         first_bf, last_bf, first_shell_bf, last_shell_bf, ncenter, kstart, katom, ktype, kprim, kshell, &
         ksumtype, Qnumber, Qstart, Qfinal, Qsbasis, Qfbasis, gccoeff, cons, gcexpo, KLMN)
   call gpu_upload_cutoff_matrix(Ycutoff, cutprim)

   call gpu_upload_method(method)
   call gpu_upload_calculated(o, co, vec, dense)
   call gpu_upload_cutoff(cutmatrix, integralCutoff, primLimit)
   call gpu_get2e(o)

   call gpu_upload_calculated(o, co, vec, dense)
   call gpu_upload_cutoff(cutmatrix, integralCutoff, primLimit)
   call gpu_upload_grad(grad, gradCutoff)
   call gpu_grad(grad)

   call gpu_cleanup()
   call gpu_shutdown()

   open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
   write(11) o, grad
   close(11)
   write(*,'(a)') 'abi_sequence: done'
end program abi_sequence
