! ref_shim.f90 -- C-callable drivers around the compiled YETI reference (oracle/_ref).
!
! TEST INFRASTRUCTURE ONLY: linked into oracle/_ref/libyetiref.so together with the
! reference's own Fortran sources (compiled where they lie by oracle/build_ref.sh). The
! reference's serial assembly and gradient drivers take assumed-shape arrays, so C cannot
! call them directly; these bind(C) wrappers give them explicit interfaces. The OpenMP
! assembly driver (explicit-size arrays) is called from Python without a wrapper.

subroutine ref_sys_linmat_serial(Kdata, Krow, Kcol, F, activeElement, nb_data, COORDS3D,    &
    &   IEN, len_IEN, nb_elem_patch, Nkv, Ukv, len_Ukv, Nijk, weight, len_weight, Jpqr,        &
    &   ELT_TYPE, len_elt, PROPS, len_PROPS, JPROPS, MATERIAL_PROPERTIES, n_mat_props, RHO,    &
    &   TENSOR, len_tensor, indDLoad, len_indDLoad, JDLType, ADLMAG, load_target_nbelem,       &
    &   load_additionalInfos, len_lai, nb_load_additionalInfos, bc_values, bc_target,          &
    &   len_bc_target, bc_target_nbelem, ind_dof_free, nb_dof_free, MCRD, NBINT, NNODE,        &
    &   nodal_dist, nb_bc, nb_load, nb_patch, nb_elem, nb_cp, nb_dof_tot, nb_n_dist,           &
    &   nb_cp_n_dist, n_mat_props_max) bind(C, name='ref_sys_linmat_serial')
    use iso_c_binding
    implicit none
    integer(c_int), intent(in) :: nb_data, len_IEN, len_Ukv, len_weight, len_elt, len_PROPS,  &
        &   len_tensor, len_indDLoad, len_lai, len_bc_target, nb_dof_free, MCRD, nb_bc,         &
        &   nb_load, nb_patch, nb_elem, nb_cp, nb_dof_tot, nb_n_dist, nb_cp_n_dist,              &
        &   n_mat_props_max
    real(c_double), intent(out) :: Kdata(nb_data), F(nb_dof_tot)
    integer(c_int), intent(out) :: Krow(nb_data), Kcol(nb_data)
    integer(c_int), intent(in) :: activeElement(nb_elem), IEN(len_IEN), nb_elem_patch(nb_patch), &
        &   Nkv(3, nb_patch), Nijk(3, nb_elem), Jpqr(3, nb_patch), JPROPS(nb_patch),             &
        &   n_mat_props(nb_patch), indDLoad(len_indDLoad), JDLType(nb_load),                     &
        &   load_target_nbelem(nb_load), nb_load_additionalInfos(nb_load), bc_target(len_bc_target), &
        &   bc_target_nbelem(nb_bc), ind_dof_free(nb_dof_tot), NBINT(nb_patch), NNODE(nb_patch)
    real(c_double), intent(in) :: COORDS3D(3, nb_cp), Ukv(len_Ukv), weight(len_weight),          &
        &   PROPS(len_PROPS), MATERIAL_PROPERTIES(n_mat_props_max, nb_patch), RHO(nb_patch),      &
        &   ADLMAG(nb_load), load_additionalInfos(len_lai), bc_values(2, nb_bc),                  &
        &   nodal_dist(nb_n_dist, nb_cp_n_dist)
    character(kind=c_char), intent(in) :: ELT_TYPE(len_elt), TENSOR(len_tensor)

    character(len=len_elt) :: elt_s
    character(len=len_tensor) :: tensor_s
    integer :: i

    interface
        subroutine sys_linmat_lindef_static(Kdata, Krow, Kcol, F, activeElement, nb_data,       &
            &   COORDS3D, IEN, nb_elem_patch, Nkv, Ukv, Nijk, weight, Jpqr, ELT_TYPE, PROPS,       &
            &   JPROPS, MATERIAL_PROPERTIES, n_mat_props, RHO, TENSOR, indDLoad, JDLType, ADLMAG,  &
            &   load_target_nbelem, load_additionalInfos, nb_load_additionalInfos, bc_values,      &
            &   nb_bc, bc_target, bc_target_nbelem, ind_dof_free, nb_dof_free, MCRD, NBINT,        &
            &   nb_load, nb_patch, nb_elem, nnode, nb_cp, nb_dof_tot, nodal_dist, nb_n_dist,       &
            &   nb_cp_n_dist, n_mat_props_max)
            integer, intent(in) :: nb_cp, nb_patch, nb_elem, nb_load, nb_bc, nb_dof_tot,        &
                &   nb_dof_free, MCRD, nb_n_dist, nb_cp_n_dist, n_mat_props_max, nb_data
            double precision, intent(in) :: COORDS3D(3, nb_cp), Ukv(:), weight(:), PROPS(:),    &
                &   MATERIAL_PROPERTIES(n_mat_props_max, nb_patch), RHO(nb_patch),               &
                &   ADLMAG(nb_load), load_additionalInfos(:), bc_values(2, nb_bc),               &
                &   nodal_dist(nb_n_dist, nb_cp_n_dist)
            integer, intent(in) :: Nkv(3, nb_patch), Jpqr(3, nb_patch), Nijk(3, nb_elem),       &
                &   NNODE(nb_patch), IEN(:), nb_elem_patch(nb_patch), JPROPS(nb_patch),          &
                &   NBINT(nb_patch), n_mat_props(nb_patch), indDLoad(:), JDLType(nb_load),       &
                &   load_target_nbelem(nb_load), nb_load_additionalInfos(:), bc_target(:),       &
                &   bc_target_nbelem(nb_bc), ind_dof_free(nb_dof_tot), activeElement(nb_elem)
            character(len=*), intent(in) :: TENSOR, ELT_TYPE
            integer, intent(out) :: Krow(nb_data), Kcol(nb_data)
            double precision, intent(out) :: Kdata(nb_data), F(nb_dof_tot)
        end subroutine
    end interface

    do i = 1, len_elt
        elt_s(i:i) = ELT_TYPE(i)
    enddo
    do i = 1, len_tensor
        tensor_s(i:i) = TENSOR(i)
    enddo

    call sys_linmat_lindef_static(Kdata, Krow, Kcol, F, activeElement, nb_data, COORDS3D, IEN,  &
        &   nb_elem_patch, Nkv, Ukv, Nijk, weight, Jpqr, elt_s, PROPS, JPROPS,                    &
        &   MATERIAL_PROPERTIES, n_mat_props, RHO, tensor_s, indDLoad, JDLType, ADLMAG,           &
        &   load_target_nbelem, load_additionalInfos, nb_load_additionalInfos, bc_values, nb_bc,  &
        &   bc_target, bc_target_nbelem, ind_dof_free, nb_dof_free, MCRD, NBINT, nb_load,         &
        &   nb_patch, nb_elem, NNODE, nb_cp, nb_dof_tot, nodal_dist, nb_n_dist, nb_cp_n_dist,     &
        &   n_mat_props_max)
end subroutine ref_sys_linmat_serial


subroutine ref_grad_linelast_work(gradWint, gradWext, computeWint, computeWext, SOL, ADJ, nadj, &
    &   activeElement, activeDir, COORDS3D, IEN, len_IEN, nb_elem_patch, Nkv, Ukv, len_Ukv,       &
    &   Nijk, weight, len_weight, Jpqr, ELT_TYPE, len_elt, PROPS, len_PROPS, JPROPS,              &
    &   MATERIAL_PROPERTIES, RHO, TENSOR, len_tensor, indDLoad, len_indDLoad, JDLType, ADLMAG,    &
    &   load_target_nbelem, load_additionalInfos, len_lai, MCRD, NBINT, nb_load, nb_patch,        &
    &   nb_elem, NNODE, nb_cp) bind(C, name='ref_grad_linelast_work')
    use iso_c_binding
    implicit none
    integer(c_int), intent(in) :: computeWint, computeWext, nadj, len_IEN, len_Ukv, len_weight,  &
        &   len_elt, len_PROPS, len_tensor, len_indDLoad, len_lai, MCRD, nb_load, nb_patch,       &
        &   nb_elem, nb_cp
    real(c_double), intent(out) :: gradWint(nadj, 3, nb_cp), gradWext(nadj, 3, nb_cp)
    real(c_double), intent(in) :: SOL(MCRD, nb_cp), ADJ(nadj, MCRD, nb_cp), COORDS3D(3, nb_cp),   &
        &   Ukv(len_Ukv), weight(len_weight), PROPS(len_PROPS), MATERIAL_PROPERTIES(2, nb_patch), &
        &   RHO(nb_patch), ADLMAG(nb_load), load_additionalInfos(len_lai)
    integer(c_int), intent(in) :: activeElement(nb_elem), activeDir(3), IEN(len_IEN),            &
        &   nb_elem_patch(nb_patch), Nkv(3, nb_patch), Nijk(3, nb_elem), Jpqr(3, nb_patch),       &
        &   JPROPS(nb_patch), indDLoad(len_indDLoad), JDLType(nb_load),                           &
        &   load_target_nbelem(nb_load), NBINT(nb_patch), NNODE(nb_patch)
    character(kind=c_char), intent(in) :: ELT_TYPE(len_elt), TENSOR(len_tensor)

    character(len=len_elt) :: elt_s
    character(len=len_tensor) :: tensor_s
    logical :: cwi, cwe
    integer :: i

    interface
        subroutine gradLinElastWork_AN(gradWint, gradWext, computeWint, computeWext, SOL, ADJ,   &
            &   nadj, activeElement, activeDir, COORDS3D, IEN, nb_elem_patch, Nkv, Ukv, Nijk,     &
            &   weight, Jpqr, ELT_TYPE, PROPS, JPROPS, MATERIAL_PROPERTIES, RHO, TENSOR,          &
            &   indDLoad, JDLType, ADLMAG, load_target_nbelem, load_additionalInfos, MCRD, NBINT, &
            &   nb_load, nb_patch, nb_elem, nnode, nb_cp)
            integer, intent(in) :: nb_cp, nb_patch, nb_elem, nb_load, MCRD, nadj
            double precision, intent(in) :: COORDS3D(3, nb_cp), Ukv(:), weight(:), PROPS(:),    &
                &   MATERIAL_PROPERTIES(2, nb_patch), RHO(nb_patch), ADLMAG(nb_load),            &
                &   load_additionalInfos(:), SOL(MCRD, nb_cp), ADJ(nadj, MCRD, nb_cp)
            integer, intent(in) :: Nkv(3, nb_patch), Jpqr(3, nb_patch), Nijk(3, nb_elem),       &
                &   NNODE(nb_patch), IEN(:), nb_elem_patch(nb_patch), JPROPS(nb_patch),          &
                &   NBINT(nb_patch), indDLoad(:), JDLType(nb_load), load_target_nbelem(nb_load), &
                &   activeElement(nb_elem), activeDir(3)
            logical, intent(in) :: computeWint, computeWext
            character(len=*), intent(in) :: TENSOR, ELT_TYPE
            double precision, intent(out) :: gradWint(nadj, 3, nb_cp), gradWext(nadj, 3, nb_cp)
        end subroutine
    end interface

    do i = 1, len_elt
        elt_s(i:i) = ELT_TYPE(i)
    enddo
    do i = 1, len_tensor
        tensor_s(i:i) = TENSOR(i)
    enddo
    cwi = computeWint /= 0
    cwe = computeWext /= 0

    call gradLinElastWork_AN(gradWint, gradWext, cwi, cwe, SOL, ADJ, nadj, activeElement,         &
        &   activeDir, COORDS3D, IEN, nb_elem_patch, Nkv, Ukv, Nijk, weight, Jpqr, elt_s, PROPS,  &
        &   JPROPS, MATERIAL_PROPERTIES, RHO, tensor_s, indDLoad, JDLType, ADLMAG,                &
        &   load_target_nbelem, load_additionalInfos, MCRD, NBINT, nb_load, nb_patch, nb_elem,    &
        &   NNODE, nb_cp)
end subroutine ref_grad_linelast_work

! reconstruction (main/reconstruction.f:22) takes bc_target(:) assumed-shape.
subroutine ref_reconstruction(sol, U, U_inv, COORDS, ind_dof_free, bc_target, len_bc_target,  &
    &   bc_target_nbelem, bc_values, nb_bc, MCRD, nb_cp, nb_dof_free)                          &
    &   bind(C, name='ref_reconstruction')
    use iso_c_binding
    implicit none
    integer(c_int), intent(in) :: len_bc_target, nb_bc, MCRD, nb_cp, nb_dof_free
    integer(c_int), intent(in) :: ind_dof_free(nb_cp * MCRD), bc_target(len_bc_target),        &
        &   bc_target_nbelem(nb_bc)
    real(c_double), intent(in) :: U_inv(nb_dof_free), COORDS(3, nb_cp), bc_values(2, nb_bc)
    real(c_double), intent(inout) :: sol(nb_cp, MCRD), U(nb_cp * MCRD)
    logical :: cpl

    interface
        subroutine reconstruction(sol, U, U_inv, COORDS, ind_dof_free, bc_target,               &
            &   bc_target_nbelem, bc_values, COUPLG_flag, nb_bc, MCRD, nb_cp, nb_dof_free)
            integer, intent(in) :: nb_cp, MCRD, nb_dof_free, nb_bc
            integer, intent(in) :: ind_dof_free(nb_cp * MCRD), bc_target(:), bc_target_nbelem(nb_bc)
            double precision, intent(in) :: U_inv(nb_dof_free), COORDS(3, nb_cp), bc_values(2, nb_bc)
            logical, intent(in) :: COUPLG_flag
            double precision, intent(out) :: sol(nb_cp, MCRD), U(nb_cp * MCRD)
        end subroutine
    end interface

    cpl = .false.
    call reconstruction(sol, U, U_inv, COORDS, ind_dof_free, bc_target, bc_target_nbelem,       &
        &   bc_values, cpl, nb_bc, MCRD, nb_cp, nb_dof_free)
end subroutine ref_reconstruction

! evalDisp (postprocessing/evaluate_disp.f:20) takes assumed-shape IEN, Ukv, weight and PROPS.
subroutine ref_evaldisp(disp, NumPatch, Xi, SOL, COORDS3D, IEN, len_IEN, nb_elem_patch, Nkv, Ukv,  &
    &   len_Ukv, Nijk, weight, len_weight, Jpqr, ELT_TYPE, len_elt, MATERIAL_PROPERTIES, TENSOR,    &
    &   len_tensor, PROPS, len_PROPS, JPROPS, NNODE, nb_patch, nb_elem, nb_cp, MCRD)               &
    &   bind(C, name='ref_evaldisp')
    use iso_c_binding
    implicit none
    integer(c_int), intent(in) :: NumPatch, len_IEN, len_Ukv, len_weight, len_elt, len_tensor,       &
        &   len_PROPS, nb_patch, nb_elem, nb_cp, MCRD
    integer(c_int), intent(in) :: IEN(len_IEN), nb_elem_patch(nb_patch), Nkv(3, nb_patch),          &
        &   Nijk(3, nb_elem), Jpqr(3, nb_patch), JPROPS(nb_patch), NNODE(nb_patch)
    real(c_double), intent(in) :: Xi(3), SOL(MCRD, nb_cp), COORDS3D(3, nb_cp), Ukv(len_Ukv),         &
        &   weight(len_weight), MATERIAL_PROPERTIES(2, nb_patch), PROPS(len_PROPS)
    character(kind=c_char), intent(in) :: ELT_TYPE(len_elt), TENSOR(len_tensor)
    real(c_double), intent(out) :: disp(3)
    character(len=len_elt) :: elt_s
    character(len=len_tensor) :: tensor_s
    integer :: i

    interface
        subroutine evalDisp(disp, NumPatch, Xi, SOL, COORDS3D, IEN, nb_elem_patch, Nkv, Ukv, Nijk,   &
            &   weight, Jpqr, ELT_TYPE, MATERIAL_PROPERTIES, TENSOR, PROPS, JPROPS, NNODE, nb_patch,  &
            &   nb_elem, nb_cp, MCRD)
            integer, intent(in) :: nb_cp, MCRD, NNODE(nb_patch), nb_patch, nb_elem, IEN(:),           &
                &   nb_elem_patch(nb_patch), JPROPS(nb_patch), Nkv(3, nb_patch), Jpqr(3, nb_patch),   &
                &   Nijk(3, nb_elem), NumPatch
            double precision, intent(in) :: COORDS3D(3, nb_cp), Ukv(:), weight(:),                  &
                &   MATERIAL_PROPERTIES(2, nb_patch), PROPS(:), SOL(MCRD, nb_cp), Xi(3)
            character(len=*), intent(in) :: TENSOR, ELT_TYPE
            double precision, intent(out) :: disp(3)
        end subroutine
    end interface

    do i = 1, len_elt
        elt_s(i:i) = ELT_TYPE(i)
    enddo
    do i = 1, len_tensor
        tensor_s(i:i) = TENSOR(i)
    enddo
    disp = 0.0d0
    call evalDisp(disp, NumPatch, Xi, SOL, COORDS3D, IEN, nb_elem_patch, Nkv, Ukv, Nijk, weight,     &
        &   Jpqr, elt_s, MATERIAL_PROPERTIES, tensor_s, PROPS, JPROPS, NNODE, nb_patch, nb_elem, nb_cp, &
        &   MCRD)
end subroutine ref_evaldisp

! build_CMASSMatrix (main/compute_CMASSMatrix.f:22) takes assumed-shape IEN, Ukv, weight, PROPS.
subroutine ref_build_cmassmatrix(Mdata, Mrow, Mcol, activeElement, nb_data, COORDS3D, IEN, len_IEN,  &
    &   nb_elem_patch, Nkv, Ukv, len_Ukv, Nijk, weight, len_weight, Jpqr, ELT_TYPE, len_elt, PROPS,   &
    &   len_PROPS, JPROPS, RHO, TENSOR, len_tensor, ind_dof_free, nb_dof_free, MCRD, NBINT, nb_patch,  &
    &   nb_elem, NNODE, nb_cp, nb_dof_tot) bind(C, name='ref_build_cmassmatrix')
    use iso_c_binding
    implicit none
    integer(c_int), intent(in) :: nb_data, len_IEN, len_Ukv, len_weight, len_elt, len_PROPS,         &
        &   len_tensor, nb_dof_free, MCRD, nb_patch, nb_elem, nb_cp, nb_dof_tot
    integer(c_int), intent(in) :: activeElement(nb_elem), IEN(len_IEN), nb_elem_patch(nb_patch),      &
        &   Nkv(3, nb_patch), Nijk(3, nb_elem), Jpqr(3, nb_patch), JPROPS(nb_patch),                  &
        &   ind_dof_free(nb_dof_tot), NBINT(nb_patch), NNODE(nb_patch)
    real(c_double), intent(in) :: COORDS3D(3, nb_cp), Ukv(len_Ukv), weight(len_weight),             &
        &   PROPS(len_PROPS), RHO(nb_patch)
    character(kind=c_char), intent(in) :: ELT_TYPE(len_elt), TENSOR(len_tensor)
    real(c_double), intent(out) :: Mdata(nb_data)
    integer(c_int), intent(out) :: Mrow(nb_data), Mcol(nb_data)
    character(len=len_elt) :: elt_s
    character(len=len_tensor) :: tensor_s
    integer :: i

    interface
        subroutine build_CMASSMatrix(Mdata, Mrow, Mcol, activeElement, nb_data, COORDS3D, IEN,       &
            &   nb_elem_patch, Nkv, Ukv, Nijk, weight, Jpqr, ELT_TYPE, PROPS, JPROPS, RHO, TENSOR,    &
            &   ind_dof_free, nb_dof_free, MCRD, NBINT, nb_patch, nb_elem, nnode, nb_cp, nb_dof_tot)
            integer, intent(in) :: nb_cp, MCRD, NNODE(nb_patch), nb_patch, NBINT(nb_patch), nb_elem,  &
                &   nb_elem_patch(nb_patch), IEN(:), JPROPS(nb_patch), nb_dof_tot, nb_dof_free,         &
                &   ind_dof_free(nb_dof_tot), nb_data, activeElement(nb_elem), Nkv(3, nb_patch),       &
                &   Jpqr(3, nb_patch), Nijk(3, nb_elem)
            double precision, intent(in) :: COORDS3D(3, nb_cp), Ukv(:), weight(:), RHO(nb_patch),    &
                &   PROPS(:)
            character(len=*), intent(in) :: TENSOR, ELT_TYPE
            double precision, intent(out) :: Mdata(nb_data)
            integer, intent(out) :: Mrow(nb_data), Mcol(nb_data)
        end subroutine
    end interface

    do i = 1, len_elt
        elt_s(i:i) = ELT_TYPE(i)
    enddo
    do i = 1, len_tensor
        tensor_s(i:i) = TENSOR(i)
    enddo
    call build_CMASSMatrix(Mdata, Mrow, Mcol, activeElement, nb_data, COORDS3D, IEN, nb_elem_patch,   &
        &   Nkv, Ukv, Nijk, weight, Jpqr, elt_s, PROPS, JPROPS, RHO, tensor_s, ind_dof_free,           &
        &   nb_dof_free, MCRD, NBINT, nb_patch, nb_elem, NNODE, nb_cp, nb_dof_tot)
end subroutine ref_build_cmassmatrix

! generateVTU (postprocessing/postproc_vtu.f:22) takes assumed-shape Ukv, weight, PROPS, IEN and
! character arguments; FieldOutput_flag arrives as 0/1 integers.
subroutine ref_generate_vtu(FILENAME, len_fn, OUTPATH, len_op, iflag, nb_refinement, SOL, COORDS3D, IEN, &
    &   len_IEN, nb_elem_patch, Nkv, Ukv, len_Ukv, Nijk, weight, len_weight, Jpqr, ELT_TYPE, len_elt,      &
    &   MATERIAL_PROPERTIES, TENSOR, len_tensor, PROPS, len_PROPS, JPROPS, NNODE, nb_patch, nb_elem,      &
    &   nb_cp, MCRD) bind(C, name='ref_generate_vtu')
    use iso_c_binding
    implicit none
    integer(c_int), intent(in) :: len_fn, len_op, len_IEN, len_Ukv, len_weight, len_elt, len_tensor,     &
        &   len_PROPS, nb_patch, nb_elem, nb_cp, MCRD
    integer(c_int), intent(in) :: iflag(3), nb_refinement(3), IEN(len_IEN), nb_elem_patch(nb_patch),    &
        &   Nkv(3, nb_patch), Nijk(3, nb_elem), Jpqr(3, nb_patch), JPROPS(nb_patch), NNODE(nb_patch)
    real(c_double), intent(in) :: SOL(MCRD, nb_cp), COORDS3D(3, nb_cp), Ukv(len_Ukv), weight(len_weight), &
        &   MATERIAL_PROPERTIES(2, nb_patch), PROPS(len_PROPS)
    character(kind=c_char), intent(in) :: FILENAME(len_fn), OUTPATH(len_op), ELT_TYPE(len_elt),          &
        &   TENSOR(len_tensor)
    character(len=len_fn) :: fn_s
    character(len=len_op) :: op_s
    character(len=len_elt) :: elt_s
    character(len=len_tensor) :: tensor_s
    logical :: flags(3)
    integer :: i

    interface
        subroutine generateVTU(FILENAME, output_path, FieldOutput_flag, nb_refinement, SOL, COORDS3D,   &
            &   IEN, nb_elem_patch, Nkv, Ukv, Nijk, weight, Jpqr, ELT_TYPE, MATERIAL_PROPERTIES, TENSOR, &
            &   PROPS, JPROPS, NNODE, nb_patch, nb_elem, nb_cp, MCRD)
            integer, intent(in) :: nb_cp, MCRD, NNODE(nb_patch), nb_patch, nb_elem, IEN(:),             &
                &   nb_elem_patch(nb_patch), JPROPS(nb_patch), Nkv(3, nb_patch), Jpqr(3, nb_patch),     &
                &   Nijk(3, nb_elem), nb_refinement(3)
            double precision, intent(in) :: COORDS3D(3, nb_cp), Ukv(:), weight(:),                    &
                &   MATERIAL_PROPERTIES(2, nb_patch), PROPS(:), SOL(MCRD, nb_cp)
            character(len=*), intent(in) :: FILENAME, output_path, TENSOR, ELT_TYPE
            logical, intent(in) :: FieldOutput_flag(3)
        end subroutine
    end interface

    do i = 1, len_fn
        fn_s(i:i) = FILENAME(i)
    enddo
    do i = 1, len_op
        op_s(i:i) = OUTPATH(i)
    enddo
    do i = 1, len_elt
        elt_s(i:i) = ELT_TYPE(i)
    enddo
    do i = 1, len_tensor
        tensor_s(i:i) = TENSOR(i)
    enddo
    do i = 1, 3
        flags(i) = iflag(i) /= 0
    enddo
    call generateVTU(fn_s, op_s, flags, nb_refinement, SOL, COORDS3D, IEN, nb_elem_patch, Nkv, Ukv,     &
        &   Nijk, weight, Jpqr, elt_s, MATERIAL_PROPERTIES, tensor_s, PROPS, JPROPS, NNODE, nb_patch,    &
        &   nb_elem, nb_cp, MCRD)
end subroutine ref_generate_vtu

! gradVibration_AN (optim/gradvibration_fullAN.f:22) takes assumed-shape Ukv, weight, PROPS, IEN.
subroutine ref_grad_vibration(gradV, VECT, VALS, nb_frq, activeElement, activeDir, COORDS3D, IEN,     &
    &   len_IEN, nb_elem_patch, Nkv, Ukv, len_Ukv, Nijk, weight, len_weight, Jpqr, ELT_TYPE, len_elt,    &
    &   PROPS, len_PROPS, JPROPS, MATERIAL_PROPERTIES, RHO, TENSOR, len_tensor, MCRD, NBINT, nb_patch,   &
    &   nb_elem, NNODE, nb_cp) bind(C, name='ref_grad_vibration')
    use iso_c_binding
    implicit none
    integer(c_int), intent(in) :: nb_frq, len_IEN, len_Ukv, len_weight, len_elt, len_PROPS, len_tensor, &
        &   MCRD, nb_patch, nb_elem, nb_cp
    integer(c_int), intent(in) :: activeElement(nb_elem), activeDir(3), IEN(len_IEN),                  &
        &   nb_elem_patch(nb_patch), Nkv(3, nb_patch), Nijk(3, nb_elem), Jpqr(3, nb_patch),              &
        &   JPROPS(nb_patch), NBINT(nb_patch), NNODE(nb_patch)
    real(c_double), intent(in) :: VECT(nb_frq, 3, nb_cp), VALS(nb_frq), COORDS3D(3, nb_cp),            &
        &   Ukv(len_Ukv), weight(len_weight), PROPS(len_PROPS), MATERIAL_PROPERTIES(2, nb_patch),        &
        &   RHO(nb_patch)
    character(kind=c_char), intent(in) :: ELT_TYPE(len_elt), TENSOR(len_tensor)
    real(c_double), intent(out) :: gradV(nb_frq, 3, nb_cp)
    character(len=len_elt) :: elt_s
    character(len=len_tensor) :: tensor_s
    integer :: i

    interface
        subroutine gradVibration_AN(gradV, VECT, VALS, nb_frq, activeElement, activeDir, COORDS3D, IEN, &
            &   nb_elem_patch, Nkv, Ukv, Nijk, weight, Jpqr, ELT_TYPE, PROPS, JPROPS,                    &
            &   MATERIAL_PROPERTIES, RHO, TENSOR, MCRD, NBINT, nb_patch, nb_elem, nnode, nb_cp)
            integer, intent(in) :: nb_cp, MCRD, NNODE(nb_patch), nb_patch, nb_elem, IEN(:),             &
                &   nb_elem_patch(nb_patch), JPROPS(nb_patch), Nkv(3, nb_patch), Jpqr(3, nb_patch),     &
                &   Nijk(3, nb_elem), NBINT(nb_patch), activeElement(nb_elem), activeDir(3), nb_frq
            double precision, intent(in) :: COORDS3D(3, nb_cp), Ukv(:), weight(:), PROPS(:),            &
                &   MATERIAL_PROPERTIES(2, nb_patch), RHO(nb_patch), VECT(nb_frq, 3, nb_cp),             &
                &   VALS(nb_frq)
            character(len=*), intent(in) :: TENSOR, ELT_TYPE
            double precision, intent(out) :: gradV(nb_frq, 3, nb_cp)
        end subroutine
    end interface

    do i = 1, len_elt
        elt_s(i:i) = ELT_TYPE(i)
    enddo
    do i = 1, len_tensor
        tensor_s(i:i) = TENSOR(i)
    enddo
    call gradVibration_AN(gradV, VECT, VALS, nb_frq, activeElement, activeDir, COORDS3D, IEN,             &
        &   nb_elem_patch, Nkv, Ukv, Nijk, weight, Jpqr, elt_s, PROPS, JPROPS, MATERIAL_PROPERTIES, RHO, &
        &   tensor_s, MCRD, NBINT, nb_patch, nb_elem, NNODE, nb_cp)
end subroutine ref_grad_vibration
