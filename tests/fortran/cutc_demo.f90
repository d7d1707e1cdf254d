! Fortran caller of cutc_int through include/cutc_module.f90.
! Mode "zero" (no argument): the reference demo's known answer -- all inputs
! zero, c = 0, m = n = o = 1 -> every output is exactly 0.  Prints
! "status <rc> maxabs <max |output|>".
program cutc_demo

    use cutc_module

    implicit none

    integer :: n_grid1, n_grid2, n_ao, n_nuc, size_bh, rc

    double precision, allocatable :: r1(:,:), wr1(:), r2(:,:), wr2(:), rn(:,:)
    double precision, allocatable :: aos_data1(:,:,:), aos_data2(:,:,:)
    double precision, allocatable :: c_bh(:,:)
    integer,          allocatable :: m_bh(:,:), n_bh(:,:), o_bh(:,:)
    double precision, allocatable :: int2_grad1_u12_ao(:,:,:,:)
    double precision, allocatable :: int_2e_ao(:,:,:,:)

    n_grid1 = 667
    n_grid2 = 6073
    n_ao = 19
    n_nuc = 3
    size_bh = 9

    allocate(r1(3,n_grid1), wr1(n_grid1))
    allocate(r2(3,n_grid2), wr2(n_grid2))
    allocate(rn(3,n_nuc))
    allocate(aos_data1(n_grid1,n_ao,4))
    allocate(aos_data2(n_grid2,n_ao,4))
    allocate(c_bh(size_bh,n_nuc), m_bh(size_bh,n_nuc), n_bh(size_bh,n_nuc), o_bh(size_bh,n_nuc))
    allocate(int2_grad1_u12_ao(n_ao,n_ao,n_grid1,3), int_2e_ao(n_ao,n_ao,n_ao,n_ao))

    r1 = 0.d0
    wr1 = 0.d0
    r2 = 0.d0
    wr2 = 0.d0
    rn = 0.d0
    aos_data1 = 0.d0
    aos_data2 = 0.d0
    c_bh = 0.d0
    m_bh = 1
    n_bh = 1
    o_bh = 1

    int2_grad1_u12_ao = 1.d0
    int_2e_ao = 1.d0

    rc = cutc_int(100, 1, 1, 32, 1, 1,                     &
                  n_grid1, n_grid2, n_ao, n_nuc, size_bh, &
                  r1, wr1, r2, wr2, rn,                   &
                  aos_data1, aos_data2,                   &
                  c_bh, m_bh, n_bh, o_bh,                 &
                  int2_grad1_u12_ao, int_2e_ao)

    print '(a,i0,a,es12.4)', 'status ', rc, ' maxabs ', &
        max(maxval(abs(int2_grad1_u12_ao)), maxval(abs(int_2e_ao)))

end program cutc_demo
