!=======================================================================
! ref_driver.f90 -- C-callable driver around the reference AdaptaHOP
! Fortran (fortran/compute_adaptahop.f90), built from the reference source
! where it lies.  TEST INFRASTRUCTURE ONLY: it is linked into
! oracle/_ref/libref_ahop.so, which tests and the CPU-baseline leg of the
! benchmark use as the checker; nothing in the product loads it.
!
! The reference exports only compute_adaptahop / sync_* / close from module
! neikdtree (compute_adaptahop.pyf:1-70).  Its stage procedures are ordinary
! global symbols, so this driver binds them by their object names to run the
! stages one at a time (tree f90:2967, density f90:682, local maxima f90:748,
! saddle list f90:2433) and to read the per-particle kNN lists
! (omp_find_nearest_parts f90:2841) while the tree is alive.
!=======================================================================
module ref_driver
   use iso_c_binding
   use fhalo_defs
   use neikdtree
   implicit none

   interface
      subroutine s_create_tree() bind(C, name="_QMneikdtreePcreate_tree_structure")
      end subroutine
      subroutine s_density() bind(C, name="_QMneikdtreePcompute_mean_density_and_np")
      end subroutine
      subroutine s_maxima() bind(C, name="_QMneikdtreePfind_local_maxima")
      end subroutine
      subroutine s_saddle() bind(C, name="_QMneikdtreePdev_compute_saddle_list")
      end subroutine
      subroutine s_knn(ipar, dist2, iparnei) bind(C, name="_QMneikdtreePomp_find_nearest_parts")
         import :: c_int, c_double
         integer(c_int) :: ipar
         real(c_double) :: dist2(0:*)
         integer(c_int) :: iparnei(*)
      end subroutine
   end interface

contains

   ! Set every module parameter the way compute_neiKDtree_mod.py does
   ! (init_adaptahop_130 :155-197, change_pos_1310 :200-210,
   ! sync_fortran :738-747) for a unit-free box of side xl.
   ! verb /= 0 turns on the reference's verbose stage prints (its own system_clock timings,
   ! on stderr); optnodes /= 0 selects its optimize_nodes create-nodes path.
   subroutine ref_configure(n, nvois, nhp, rho_thr, fdg, alph, fdgeps, nmemb, nlev, nthreads, xl, &
         verb, optnodes) bind(C, name="ref_configure")
      integer(c_int64_t), value :: n
      integer(c_int), value :: nvois, nhp, nmemb, nlev, nthreads, verb, optnodes
      real(c_double), value :: rho_thr, fdg, alph, fdgeps, xl
      real(c_double) :: shift(3), eps
      integer(c_int) :: n4
      n4 = int(n, c_int)
      shift = 0d0
      call sync_from_init_adaptahop(n, nmemb, nmemb, 0.7d0, 0.7d0, 0.3d0, 0.3d0, 1d0, 1d0, &
           0.7d0, 70d0, xl, xl, xl, xl, xl, xl*0.5d0, xl*0.5d0, xl*0.5d0, 70d0, 1d0, shift)
      eps = fdgeps*xl/dble(n)**(1d0/3d0)
      call sync_from_change_pos(n, n4, eps, fdgeps, xl, xl)
      call sync_others(verb /= 0, .false., n4, nthreads, rho_thr, 1d0/dble(n), 0d0, nhp, nvois, &
           fdg, alph, 'MSM', nlev, optnodes /= 0)
   end subroutine ref_configure

   ! Full public entry point, exactly as the Python bridge calls it.
   subroutine ref_run_full(n, pos_in, mass_in) bind(C, name="ref_run_full")
      integer(c_int64_t), value :: n
      real(c_double), intent(in) :: pos_in(3, n), mass_in(n)
      call compute_adaptahop(pos_in, mass_in)
   end subroutine ref_run_full

   ! Stage-by-stage run: copy-in (f90:501-518), tree, density, maxima.
   subroutine ref_stage_begin(n, pos_in, mass_in) bind(C, name="ref_stage_begin")
      integer(c_int64_t), value :: n
      real(c_double), intent(in) :: pos_in(3, n), mass_in(n)
      call close()
      allocate(liste_parts(1:nusedpart))
      allocate(pos(1:3, 1:npart))
      pos(:, :) = pos_in(:, :)
      allocate(mass(npart))
      mass(:) = mass_in(:)
   end subroutine ref_stage_begin

   subroutine ref_stage_tree() bind(C, name="ref_stage_tree")
      call s_create_tree()
   end subroutine

   subroutine ref_stage_density() bind(C, name="ref_stage_density")
      call s_density()
   end subroutine

   subroutine ref_stage_maxima() bind(C, name="ref_stage_maxima")
      call s_maxima()
   end subroutine

   subroutine ref_stage_saddle() bind(C, name="ref_stage_saddle")
      call s_saddle()
   end subroutine

   ! kNN of particle ipar (1-based) while the tree is alive.
   subroutine ref_knn(ipar, dist2, iparnei) bind(C, name="ref_knn")
      integer(c_int), value :: ipar
      real(c_double) :: dist2(0:*)
      integer(c_int) :: iparnei(*)
      integer(c_int) :: ip
      ip = ipar
      call s_knn(ip, dist2, iparnei)
   end subroutine

   subroutine ref_close() bind(C, name="ref_close")
      call close()
      if (allocated(pos)) deallocate(pos)
      if (allocated(mass)) deallocate(mass)
   end subroutine

   integer(c_int) function ref_ngroups() bind(C, name="ref_ngroups")
      ref_ngroups = ngroups
   end function

   integer(c_int) function ref_nnodes() bind(C, name="ref_nnodes")
      ref_nnodes = nnodes
   end function

   integer(c_int) function ref_nactive() bind(C, name="ref_nactive")
      ref_nactive = nactive_parts
   end function

   subroutine ref_get_density(out, n) bind(C, name="ref_get_density")
      integer(c_int64_t), value :: n
      real(c_double) :: out(n)
      out(1:n) = density(1:n)
   end subroutine

   subroutine ref_get_liste_parts(out, n) bind(C, name="ref_get_liste_parts")
      integer(c_int64_t), value :: n
      integer(c_int) :: out(n)
      out(1:n) = liste_parts(1:n)
   end subroutine

   subroutine ref_get_densityg(out, ng) bind(C, name="ref_get_densityg")
      integer(c_int), value :: ng
      real(c_double) :: out(ng)
      out(1:ng) = densityg(1:ng)
   end subroutine

   ! Saddle CSR after dev_compute_saddle_list (f90:2433-2813).  Offsets are
   ! 1-based as in the reference; counts come from offsets.
   integer(c_int) function ref_saddle_total() bind(C, name="ref_saddle_total")
      ref_saddle_total = dev_group_offset(ngroups + 1) - 1
   end function

   subroutine ref_get_saddle(offset, isad, rho) bind(C, name="ref_get_saddle")
      integer(c_int) :: offset(*), isad(*)
      real(c_double) :: rho(*)
      integer :: ne
      ne = dev_group_offset(ngroups + 1) - 1
      offset(1:ngroups + 1) = dev_group_offset(1:ngroups + 1)
      if (ne > 0) then
         isad(1:ne) = dev_isad_all(1:ne)
         rho(1:ne) = dev_rho_saddle_all(1:ne)
      endif
   end subroutine

   subroutine ref_get_tables(rt, it) bind(C, name="ref_get_tables")
      real(c_double) :: rt(8, *)
      integer(c_int) :: it(6, *)
      if (nnodes > 0) then
         rt(:, 1:nnodes) = real_table(:, 1:nnodes)
         it(:, 1:nnodes) = integer_table(:, 1:nnodes)
      endif
   end subroutine

end module ref_driver
