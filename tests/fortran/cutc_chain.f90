! Fortran caller of the whole chain through include/cutc_module.f90:
! AOs on both grids (cutc_aos) -> cutc_int -> cutc_ao_to_mo -> cutc_no.
! Reads its inputs from the stream file given as the first argument and
! writes the status codes and outputs to the second (layouts in the test
! that drives it, tests/test_gpu_parity.py).
program cutc_chain

    use cutc_module

    implicit none

    integer :: hdr(9), n_nuc, n_ao, n1, n2, size_bh, prim_max, n_mo, ne_a, ne_b, st(5), u
    character(len=4096) :: fin, fout
    double precision, allocatable :: r1(:,:), wr1(:), r2(:,:), wr2(:), rn(:,:)
    integer, allocatable :: ao_nucl(:), ao_power(:,:), ao_prim_num(:)
    double precision, allocatable :: ao_expo(:,:), ao_coef(:,:), c_bh(:,:), cl(:,:), cr(:,:)
    integer, allocatable :: m_bh(:,:), n_bh(:,:), o_bh(:,:)
    double precision, allocatable :: aos1(:,:,:), aos2(:,:,:), x(:,:,:,:), i2e(:,:,:,:)
    double precision, allocatable :: g(:,:,:,:), mol(:,:), mor(:,:), no2(:,:,:,:), no1(:,:)
    double precision :: no0

    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
    read(u) hdr
    n_nuc = hdr(1); n_ao = hdr(2); n1 = hdr(3); n2 = hdr(4); size_bh = hdr(5)
    prim_max = hdr(6); n_mo = hdr(7); ne_a = hdr(8); ne_b = hdr(9)
    allocate(r1(3,n1), wr1(n1), r2(3,n2), wr2(n2), rn(3,n_nuc))
    allocate(ao_nucl(n_ao), ao_power(n_ao,3), ao_prim_num(n_ao))
    allocate(ao_expo(prim_max,n_ao), ao_coef(prim_max,n_ao))
    allocate(c_bh(size_bh,n_nuc), m_bh(size_bh,n_nuc), n_bh(size_bh,n_nuc), o_bh(size_bh,n_nuc))
    allocate(cl(n_ao,n_mo), cr(n_ao,n_mo))
    read(u) r1, wr1, r2, wr2, rn
    read(u) ao_nucl, ao_power, ao_prim_num
    read(u) ao_expo, ao_coef
    read(u) c_bh
    read(u) m_bh, n_bh, o_bh
    read(u) cl, cr
    close(u)

    allocate(aos1(n1,n_ao,4), aos2(n2,n_ao,4))
    allocate(x(n_ao,n_ao,n1,3), i2e(n_ao,n_ao,n_ao,n_ao))
    allocate(g(n1,3,n_mo,n_mo), mol(n1,n_mo), mor(n1,n_mo))
    allocate(no2(n_mo,n_mo,n_mo,n_mo), no1(n_mo,n_mo))

    st(1) = cutc_aos(n1, n_nuc, n_ao, prim_max, r1, rn, ao_nucl, ao_power, ao_prim_num, &
                     ao_expo, ao_coef, aos1)
    st(2) = cutc_aos(n2, n_nuc, n_ao, prim_max, r2, rn, ao_nucl, ao_power, ao_prim_num, &
                     ao_expo, ao_coef, aos2)
    st(3) = cutc_int(100, 1, 1, 32, 1, 1, n1, n2, n_ao, n_nuc, size_bh, &
                     r1, wr1, r2, wr2, rn, aos1, aos2, c_bh, m_bh, n_bh, o_bh, x, i2e)
    st(4) = cutc_ao_to_mo(n1, n_ao, n_mo, cl, cr, x, g)
    mol = matmul(aos1(:,:,1), cl)
    mor = matmul(aos1(:,:,1), cr)
    st(5) = cutc_no(n1, n_mo, ne_a, ne_b, wr1, mol, mor, g, no2, no1, no0)

    open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(u) st
    write(u) i2e, g, no2, no1, no0
    close(u)
    print '(a,5i4)', 'status', st

end program cutc_chain
