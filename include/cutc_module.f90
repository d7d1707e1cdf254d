! Fortran binding of libcutcint's cutc_int, cutc_no, cutc_ao_to_mo and cutc_aos (bind(C)).
!
! Array shapes follow the reference binding src/cutc_module.f90:23-35.  The
! argument list follows the C definition src/cutc_int.c:15-20, which takes six
! launch-geometry integers (grid x, y, z and block x, y, z); the reference
! module declares only two (src/cutc_module.f90:10-18), which shifts every
! later argument.  The function returns the C status (0, or a negative code).
module cutc_module

  use, intrinsic :: iso_c_binding

  implicit none

  interface

    integer(c_int) function cutc_int(nxBlocks, nyBlocks, nzBlocks,          &
                                     blockxSize, blockySize, blockzSize,    &
                                     n_grid1, n_grid2, n_ao, n_nuc, size_bh, &
                                     r1, wr1, r2, wr2, rn,                   &
                                     aos_data1, aos_data2,                   &
                                     c_bh, m_bh, n_bh, o_bh,                 &
                                     int2_grad1_u12_ao, int_2e_ao) bind(C, name = "cutc_int")

      import c_int, c_double
      integer(c_int), intent(in), value :: nxBlocks, nyBlocks, nzBlocks
      integer(c_int), intent(in), value :: blockxSize, blockySize, blockzSize
      integer(c_int), intent(in), value :: n_grid1, n_grid2
      integer(c_int), intent(in), value :: n_ao
      integer(c_int), intent(in), value :: n_nuc
      integer(c_int), intent(in), value :: size_bh
      real(c_double), intent(in)        :: r1(3,n_grid1)
      real(c_double), intent(in)        :: wr1(n_grid1)
      real(c_double), intent(in)        :: r2(3,n_grid2)
      real(c_double), intent(in)        :: wr2(n_grid2)
      real(c_double), intent(in)        :: rn(3,n_nuc)
      real(c_double), intent(in)        :: aos_data1(n_grid1,n_ao,4)
      real(c_double), intent(in)        :: aos_data2(n_grid2,n_ao,4)
      real(c_double), intent(in)        :: c_bh(size_bh,n_nuc)
      integer(c_int), intent(in)        :: m_bh(size_bh,n_nuc)
      integer(c_int), intent(in)        :: n_bh(size_bh,n_nuc)
      integer(c_int), intent(in)        :: o_bh(size_bh,n_nuc)
      real(c_double), intent(out)       :: int2_grad1_u12_ao(n_ao,n_ao,n_grid1,3)
      real(c_double), intent(out)       :: int_2e_ao(n_ao,n_ao,n_ao,n_ao)

    end function cutc_int

    ! Normal ordering (src/cutc_no.c:14-16; the reference module has no
    ! binding for it).  Returns the C status.
    integer(c_int) function cutc_no(n_grid1, n_mo, ne_a, ne_b,              &
                                    wr1, mos_l_in_r, mos_r_in_r,            &
                                    int2_grad1_u12, no_2e, no_1e, no_0e)    &
                                    bind(C, name = "cutc_no")

      import c_int, c_double
      integer(c_int), intent(in), value :: n_grid1, n_mo, ne_a, ne_b
      real(c_double), intent(in)        :: wr1(n_grid1)
      real(c_double), intent(in)        :: mos_l_in_r(n_grid1,n_mo)
      real(c_double), intent(in)        :: mos_r_in_r(n_grid1,n_mo)
      real(c_double), intent(in)        :: int2_grad1_u12(n_grid1,3,n_mo,n_mo)
      real(c_double), intent(out)       :: no_2e(n_mo,n_mo,n_mo,n_mo)
      real(c_double), intent(out)       :: no_1e(n_mo,n_mo)
      real(c_double), intent(out)       :: no_0e

    end function cutc_no

    ! AO -> MO transform of int2_grad1_u12 between cutc_int and cutc_no (no
    ! counterpart in the reference): int2_grad1_u12(g,x,i,j) =
    ! sum_{k,l} mo_l_coef(k,i) int2_grad1_u12_ao(k,l,g,x) mo_r_coef(l,j).
    integer(c_int) function cutc_ao_to_mo(n_grid1, n_ao, n_mo,              &
                                          mo_l_coef, mo_r_coef,             &
                                          int2_grad1_u12_ao, int2_grad1_u12)&
                                          bind(C, name = "cutc_ao_to_mo")

      import c_int, c_double
      integer(c_int), intent(in), value :: n_grid1, n_ao, n_mo
      real(c_double), intent(in)        :: mo_l_coef(n_ao,n_mo)
      real(c_double), intent(in)        :: mo_r_coef(n_ao,n_mo)
      real(c_double), intent(in)        :: int2_grad1_u12_ao(n_ao,n_ao,n_grid1,3)
      real(c_double), intent(out)       :: int2_grad1_u12(n_grid1,3,n_mo,n_mo)

    end function cutc_ao_to_mo

    ! AO values and gradients on a grid, aos_data(n_grid,n_ao,4) as cutc_int
    ! reads it (no counterpart in the reference).  ao_nucl holds 1-based
    ! nucleus numbers.
    integer(c_int) function cutc_aos(n_grid, n_nuc, n_ao, prim_max, r, rn,  &
                                     ao_nucl, ao_power, ao_prim_num,        &
                                     ao_expo, ao_coef, aos_data)            &
                                     bind(C, name = "cutc_aos")

      import c_int, c_double
      integer(c_int), intent(in), value :: n_grid, n_nuc, n_ao, prim_max
      real(c_double), intent(in)        :: r(3,n_grid), rn(3,n_nuc)
      integer(c_int), intent(in)        :: ao_nucl(n_ao), ao_power(n_ao,3)
      integer(c_int), intent(in)        :: ao_prim_num(n_ao)
      real(c_double), intent(in)        :: ao_expo(prim_max,n_ao)
      real(c_double), intent(in)        :: ao_coef(prim_max,n_ao)
      real(c_double), intent(out)       :: aos_data(n_grid,n_ao,4)

    end function cutc_aos

  end interface

end module cutc_module
