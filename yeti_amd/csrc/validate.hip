// validate.hip -- host-side checks of a yeti_system_inputs before any device work.
//
// The assembly and matrix-free set-ups gather control points, weights and knots through the
// ids the caller passes (IEN, Nijk, load and boundary-condition targets). YETI's Fortran trusts
// them (build_system_elemStorage.f90:156-293 indexes COORDS3D(:, IEN(...)) directly); here every
// id is range-checked first, so a malformed model returns YETI_ERR_INVALID_ARGUMENT instead of
// reaching a kernel or a host table with an out-of-range index. Runs on the host only: a
// rejected model never touches the device.
#include <string>
#include <vector>

#include "common.hpp"

namespace yeti {

static bool solid_type(const std::string &t) { return t == "U1" || t == "U0"; }

// Element-type names of ELT_TYPE, one per patch ("U1U1U0" -> U1, U1, U0); false when fewer
// names than patches are listed.
static bool split_types(const yeti_system_inputs *in, std::vector<std::string> &out) {
    out.clear();
    if (!in->ELT_TYPE || in->len_ELT_TYPE <= 0) return false;
    const std::string s(in->ELT_TYPE, in->ELT_TYPE + in->len_ELT_TYPE);
    size_t k = 0;
    while (k < s.size()) {
        if (s[k] != 'U') {
            ++k;
            continue;
        }
        size_t j = k + 1;
        while (j < s.size() && s[j] != 'U' && s[j] != ' ' && s[j] != '\0') ++j;
        out.push_back(s.substr(k, j - k));
        k = j;
    }
    return (int)out.size() >= in->nb_patch;
}

int validate_system_inputs(const yeti_system_inputs *in, bool require_solid) {
    if (!in) return set_error(YETI_ERR_INVALID_ARGUMENT, "null inputs");
    if (in->MCRD != 3) return set_error(YETI_ERR_INVALID_ARGUMENT, "MCRD=%d, only 3D solids are supported", in->MCRD);
    if (in->nb_patch < 1 || in->nb_cp < 1 || in->nb_elem < 0 || in->nb_load < 0 || in->nb_bc < 0 ||
        in->nb_dof_tot < 0)
        return set_error(YETI_ERR_INVALID_ARGUMENT, "empty model or negative count");
    if ((int64_t)in->nb_dof_tot != 3LL * in->nb_cp)
        return set_error(YETI_ERR_INVALID_ARGUMENT, "nb_dof_tot=%d with %d control points", in->nb_dof_tot, in->nb_cp);
    if (in->n_mat_props_max < 2) return set_error(YETI_ERR_INVALID_ARGUMENT, "n_mat_props_max=%d < 2", in->n_mat_props_max);
    if (!in->COORDS3D || !in->IEN || !in->nb_elem_patch || !in->Nkv || !in->Ukv || !in->Jpqr || !in->NBINT ||
        !in->NNODE || !in->MATERIAL_PROPERTIES || !in->RHO)
        return set_error(YETI_ERR_INVALID_ARGUMENT, "null model array");
    if (in->nb_elem > 0 && (!in->activeElement || !in->Nijk || !in->weight))
        return set_error(YETI_ERR_INVALID_ARGUMENT, "null element array");
    std::vector<std::string> types;
    if (!split_types(in, types)) return set_error(YETI_ERR_INVALID_ARGUMENT, "ELT_TYPE names fewer types than patches");

    // per-patch sizes and the lengths of the concatenated arrays
    int64_t nelem = 0, nkv_sum = 0, ien_len = 0;
    for (int ip = 0; ip < in->nb_patch; ++ip) {
        const int ne = in->nb_elem_patch[ip];
        if (ne < 0) return set_error(YETI_ERR_INVALID_ARGUMENT, "patch %d: %d elements", ip + 1, ne);
        if (in->NNODE[ip] < 1) return set_error(YETI_ERR_INVALID_ARGUMENT, "patch %d: NNODE=%d", ip + 1, in->NNODE[ip]);
        if (in->NBINT[ip] < 1) return set_error(YETI_ERR_INVALID_ARGUMENT, "patch %d: NBINT=%d", ip + 1, in->NBINT[ip]);
        for (int d = 0; d < 3; ++d) {
            const int p = in->Jpqr[3 * ip + d], nk = in->Nkv[3 * ip + d];
            if (p < 0 || nk < 2 * (p + 1))
                return set_error(YETI_ERR_INVALID_ARGUMENT, "patch %d: degree %d with %d knots in direction %d", ip + 1,
                                 p, nk, d + 1);
            nkv_sum += nk;
        }
        nelem += ne;
        ien_len += (int64_t)in->NNODE[ip] * ne;
        if (require_solid && !solid_type(types[ip]))
            return set_error(YETI_ERR_UNSUPPORTED_ELEMENT, "patch %d: element type %s is not supported", ip + 1,
                             types[ip].c_str());
    }
    if (nelem != in->nb_elem)
        return set_error(YETI_ERR_INVALID_ARGUMENT, "nb_elem_patch sums to %lld, nb_elem=%d", (long long)nelem, in->nb_elem);
    if (in->len_Ukv < nkv_sum)
        return set_error(YETI_ERR_INVALID_ARGUMENT, "Ukv holds %lld knots, Nkv names %lld", (long long)in->len_Ukv,
                         (long long)nkv_sum);
    if (in->len_IEN < ien_len)
        return set_error(YETI_ERR_INVALID_ARGUMENT, "IEN holds %lld ids, the patches need %lld", (long long)in->len_IEN,
                         (long long)ien_len);
    if (in->len_weight < ien_len)
        return set_error(YETI_ERR_INVALID_ARGUMENT, "weight holds %lld values, the patches need %lld",
                         (long long)in->len_weight, (long long)ien_len);

    // connectivity and knot spans, element by element
    int64_t ieoff = 0, eloff = 0, kv = 0;
    for (int ip = 0; ip < in->nb_patch; ++ip) {
        const int ne = in->nb_elem_patch[ip], nn = in->NNODE[ip];
        for (int64_t k = 0; k < (int64_t)nn * ne; ++k) {
            const int32_t id = in->IEN[ieoff + k];
            if (id < 1 || id > in->nb_cp)
                return set_error(YETI_ERR_INVALID_ARGUMENT, "patch %d, element %lld: IEN entry %d not in [1, %d]", ip + 1,
                                 (long long)(k / nn + 1), id, in->nb_cp);
        }
        if (solid_type(types[ip])) {
            const double *U = in->Ukv + kv;
            for (int d = 0; d < 3; ++d) {
                const int p = in->Jpqr[3 * ip + d], nk = in->Nkv[3 * ip + d];
                for (int i = 1; i < nk; ++i)
                    if (!(U[i] >= U[i - 1]))
                        return set_error(YETI_ERR_INVALID_ARGUMENT, "patch %d: knot vector %d decreases at knot %d",
                                         ip + 1, d + 1, i + 1);
                // a 1-based span ni reads knots ni-p+1 .. ni+p (dersbasisfuns.f90:60-62); the
                // element lies in [U(ni), U(ni+1)] with ni in [p+1, nk-p-1]
                for (int e = 0; e < ne; ++e) {
                    const int32_t s = in->Nijk[3 * (eloff + e) + d];
                    if (s < p + 1 || s > nk - p - 1)
                        return set_error(YETI_ERR_INVALID_ARGUMENT,
                                         "patch %d, element %d: span %d in direction %d not in [%d, %d]", ip + 1, e + 1,
                                         s, d + 1, p + 1, nk - p - 1);
                }
                U += nk;
            }
        }
        for (int d = 0; d < 3; ++d) kv += in->Nkv[3 * ip + d];
        ieoff += (int64_t)nn * ne;
        eloff += ne;
    }

    // loads: element loads name elements, concentrated loads name control points
    if (in->nb_load > 0) {
        if (!in->JDLType || !in->ADLMAG || !in->load_target_nbelem || !in->nb_load_additionalInfos)
            return set_error(YETI_ERR_INVALID_ARGUMENT, "null load array");
        int64_t ntarget = 0;
        for (int il = 0; il < in->nb_load; ++il) {
            if (in->load_target_nbelem[il] < 0 || in->nb_load_additionalInfos[il] < 0)
                return set_error(YETI_ERR_INVALID_ARGUMENT, "load %d: negative count", il + 1);
            ntarget += in->load_target_nbelem[il];
        }
        if (ntarget > 0 && (!in->indDLoad || in->len_indDLoad < ntarget))
            return set_error(YETI_ERR_INVALID_ARGUMENT, "indDLoad holds %lld targets, the loads name %lld",
                             (long long)in->len_indDLoad, (long long)ntarget);
        int64_t off = 0;
        for (int il = 0; il < in->nb_load; ++il) {
            const int t = in->JDLType[il];
            const bool elem_load = (t > 9 && t < 100) || t == 101 || t == 102;
            const bool conc = t / 10 < 1;
            const int32_t hi = elem_load ? in->nb_elem : in->nb_cp;
            for (int32_t k = 0; k < in->load_target_nbelem[il]; ++k) {
                const int32_t id = in->indDLoad[off + k];
                if ((elem_load || conc) && (id < 1 || id > hi))
                    return set_error(YETI_ERR_INVALID_ARGUMENT, "load %d: target %d not in [1, %d]", il + 1, id, hi);
            }
            off += in->load_target_nbelem[il];
        }
    }
    // boundary conditions name control points
    if (in->nb_bc > 0) {
        if (!in->bc_target_nbelem || !in->bc_values) return set_error(YETI_ERR_INVALID_ARGUMENT, "null boundary-condition array");
        int64_t ntarget = 0;
        for (int ib = 0; ib < in->nb_bc; ++ib) {
            if (in->bc_target_nbelem[ib] < 0) return set_error(YETI_ERR_INVALID_ARGUMENT, "condition %d: negative count", ib + 1);
            ntarget += in->bc_target_nbelem[ib];
        }
        if (ntarget > 0 && (!in->bc_target || in->len_bc_target < ntarget))
            return set_error(YETI_ERR_INVALID_ARGUMENT, "bc_target holds %lld ids, the conditions name %lld",
                             (long long)in->len_bc_target, (long long)ntarget);
        for (int64_t k = 0; k < ntarget; ++k)
            if (in->bc_target[k] < 1 || in->bc_target[k] > in->nb_cp)
                return set_error(YETI_ERR_INVALID_ARGUMENT, "bc_target entry %d not in [1, %d]", in->bc_target[k], in->nb_cp);
    }
    if (in->ind_dof_free)
        for (int32_t k = 0; k < in->nb_dof_tot; ++k)
            if (in->ind_dof_free[k] < 0 || in->ind_dof_free[k] > in->nb_dof_tot)
                return set_error(YETI_ERR_INVALID_ARGUMENT, "ind_dof_free entry %d not in [0, %d]", in->ind_dof_free[k],
                                 in->nb_dof_tot);
    if (in->nb_dof_free < 0 || in->nb_dof_free > in->nb_dof_tot)
        return set_error(YETI_ERR_INVALID_ARGUMENT, "nb_dof_free=%d with %d dofs", in->nb_dof_free, in->nb_dof_tot);
    return 0;
}

}  // namespace yeti
