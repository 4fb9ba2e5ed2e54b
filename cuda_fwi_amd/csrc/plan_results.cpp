// plan_results.cpp -- results read back to the host: traces, fields, snapshots, the
// forward store and the gradients.
#include "plan.h"

namespace psvplan {

template <typename T>
psvfwi_status Plan<T>::get_snapshot(int shot, int k, double* out) const
{
    if (shot < 0 || shot >= nslots_ || !out) return fail(PSVFWI_E_ARG, "get_snapshot: bad shot");
    const int s = slot_of_shot_[(size_t)shot];
    const int n = nsnap_slot(s);
    if (k < 0 || k >= n) return fail(PSVFWI_E_ARG, "get_snapshot: no such snapshot");
    std::vector<double> h((size_t)g_.nzt * pitch_);
    PSV_TRY(download(snaps_.p + ((size_t)k * nslots_ + s) * fst_ + org_, h.size(), h.data()));
    for (int iz = 0; iz < g_.nzt; iz++)
        std::memcpy(out + (size_t)iz * g_.nxt, h.data() + (size_t)iz * pitch_, (size_t)g_.nxt * sizeof(double));
    return PSVFWI_OK;
}

template <typename T>
psvfwi_status Plan<T>::snapshot_count(int shot, int* n) const
{
    if (shot < 0 || shot >= nslots_ || !n) return fail(PSVFWI_E_ARG, "snapshot_count: bad shot");
    *n = snaps_taken_ ? nsnap_slot(slot_of_shot_[(size_t)shot]) : 0;
    return PSVFWI_OK;
}

template <typename T>
psvfwi_status Plan<T>::get_records(int shot, double* out) const
{
    if (shot < 0 || shot >= nslots_ || !out) return fail(PSVFWI_E_ARG, "get_records: bad shot");
    if (!fwd_done_) return fail(PSVFWI_E_STATE, "get_records: run forward first");
    const int s = slot_of_shot_[(size_t)shot];
    const size_t len = shots_[(size_t)shot].rx.size() * (size_t)shots_[(size_t)shot].nt;
    return download(rec_.p + rec_out_off_h_[(size_t)s], len, out);
}

template <typename T>
psvfwi_status Plan<T>::get_field(int shot, int field, double* out) const
{
    if (shot < 0 || shot >= nslots_ || !out) return fail(PSVFWI_E_ARG, "get_field: bad shot");
    if (!ran_) return fail(PSVFWI_E_STATE, "get_field: nothing has run");
    const int s = slot_of_shot_[(size_t)shot];
    const int nzt = g_.nzt, nxt = g_.nxt;
    const State<T>& cur = (parity_ && ((ntmax_ - nt_slot_[(size_t)s]) & 1)) ? main2_ : main_;
    if (field >= 0 && field <= 7) {
        const T* base = nullptr;
        if (field <= 4) base = cur.f[field].p;
        else if (field == 5) base = We_.p;
        else if (field == 6) base = ux_.p;
        else base = uz_.p;
        if (!base) return fail(PSVFWI_E_STATE, "get_field: field not allocated");
        std::vector<double> h((size_t)nzt * pitch_);
        PSV_TRY(download(base + (size_t)s * fst_ + org_, h.size(), h.data()));
        for (int iz = 0; iz < nzt; iz++)
            std::memcpy(out + (size_t)iz * nxt, h.data() + (size_t)iz * pitch_, (size_t)nxt * sizeof(double));
        return PSVFWI_OK;
    }
    if (field >= 8 && field <= 15) {
        if (g_.npml <= 0) return fail(PSVFWI_E_ARG, "get_field: no CPML");
        const int k = field - 8;
        // 8 mem_vx_x, 9 mem_vz_x (stress side, x), 10 mem_sxx_x, 11 mem_szx_x (velocity side, x)
        // 12 mem_vx_z, 13 mem_vz_z (stress side, z), 14 mem_szz_z, 15 mem_szx_z (velocity side, z)
        const T* base = nullptr;
        bool xtype = k < 4;
        switch (k) {
        case 0: base = cur.ms[0].p; break;
        case 1: base = cur.ms[1].p; break;
        case 2: base = main_.mv[0].p; break;
        case 3: base = main_.mv[1].p; break;
        case 4: base = cur.ms[2].p; break;
        case 5: base = cur.ms[3].p; break;
        case 6: base = main_.mv[2].p; break;
        default: base = main_.mv[3].p; break;
        }
        if (xtype) return download(base + (size_t)s * xst_, (size_t)nzt * tp_, out);
        std::vector<double> h((size_t)tp_ * pitch_);
        PSV_TRY(download(base + (size_t)s * zst_, h.size(), h.data()));
        for (int r = 0; r < tp_; r++)
            std::memcpy(out + (size_t)r * nxt, h.data() + (size_t)r * pitch_, (size_t)nxt * sizeof(double));
        return PSVFWI_OK;
    }
    return fail(PSVFWI_E_ARG, "get_field: unknown field");
}

template <typename T>
psvfwi_status Plan<T>::get_store(int shot, double* out) const
{
    if (shot < 0 || shot >= nslots_ || !out) return fail(PSVFWI_E_ARG, "get_store: bad shot");
    if (!store_loaded_ || store_mode_ != PSVFWI_STORE_FULL)
        return fail(PSVFWI_E_STATE, "get_store: no full forward store held");
    const int s = slot_of_shot_[(size_t)shot];
    const int nft = 1 + (shots_[(size_t)shot].nt - 1) / w_.dt;
    const size_t plane = (size_t)nfz_ * nfx_;
    if (sst_ < (size_t)nft * 5 * plane || store_.n < (size_t)nslots_ * sst_)
        return fail(PSVFWI_E_STATE, "get_store: the store does not match the current window and shots");
    std::vector<double> h(sst_);
    PSV_TRY(download(store_.p + (size_t)s * sst_, sst_, h.data()));
    // the window loops run over [z1, z2) x [x1, x2): when an extent is a
    // multiple of its stride the last row / column of the store is never
    // written (CUDA_week2/forward_kernel_PSV_GPU.cu:803-821) and reads 0
    const int nzw = (w_.z2 - w_.z1 + w_.dz - 1) / w_.dz, nxw = (w_.x2 - w_.x1 + w_.dx - 1) / w_.dx;
    for (int f = 0; f < 5; f++)
        for (int t = 0; t < nft; t++) {
            double* o = out + ((size_t)f * nft + t) * plane;
            const double* src = h.data() + ((size_t)t * 5 + f) * plane;
            for (int zf = 0; zf < nfz_; zf++)
                for (int xf = 0; xf < nfx_; xf++) {
                    const size_t k = (size_t)zf * nfx_ + xf;
                    o[k] = (zf < nzw && xf < nxw) ? src[k] : 0.0;
                }
        }
    return PSVFWI_OK;
}

template <typename T>
psvfwi_status Plan<T>::get_gradient(int shot, double* gl, double* gm, double* gr) const
{
    if (!adj_grad_done_) return fail(PSVFWI_E_STATE, "get_gradient: run adjoint with PSVFWI_GRADIENT first");
    if (shot < -1 || shot >= nslots_ || !gl || !gm || !gr) return fail(PSVFWI_E_ARG, "get_gradient: bad shot");
    const size_t plane = (size_t)nfz_ * nfx_;
    if (gst_ < 3 * plane || grad_.n < (size_t)nslots_ * gst_)
        return fail(PSVFWI_E_STATE, "get_gradient: the gradient does not match the current window and shots");
    std::vector<double> h(3 * plane);
    if (shot >= 0) {
        const int s = slot_of_shot_[(size_t)shot];
        PSV_TRY(download(grad_.p + (size_t)s * gst_, 3 * plane, h.data()));
    } else {
        DBuf<double> sum;
        DBuf<int> order;
        PSV_HIP(sum.alloc(3 * plane));
        PSV_HIP(order.alloc((size_t)nslots_));
        PSV_HIP(hipMemcpy(order.p, slot_of_shot_.data(), (size_t)nslots_ * sizeof(int), hipMemcpyHostToDevice));
        PSV_HIP(psv::launch_grad_sum<T>(grad_.p, gst_, order.p, nslots_, 3 * plane, sum.p, stream_));
        PSV_HIP(hipMemcpyAsync(h.data(), sum.p, 3 * plane * sizeof(double), hipMemcpyDeviceToHost, stream_));
        PSV_HIP(hipStreamSynchronize(stream_));
    }
    std::memcpy(gl, h.data(), plane * sizeof(double));
    std::memcpy(gm, h.data() + plane, plane * sizeof(double));
    std::memcpy(gr, h.data() + 2 * plane, plane * sizeof(double));
    return PSVFWI_OK;
}

template <typename T>
psvfwi_status Plan<T>::download(const T* src, size_t n, double* out) const
{
    if (n == 0) return PSVFWI_OK;
    std::vector<T> h(n);
    PSV_HIP(hipMemcpyAsync(h.data(), src, n * sizeof(T), hipMemcpyDeviceToHost, stream_));
    PSV_HIP(hipStreamSynchronize(stream_));
    for (size_t i = 0; i < n; i++) out[i] = (double)h[i];
    return PSVFWI_OK;
}

#define PSV_INSTANTIATE(T) \
    template psvfwi_status Plan<T>::get_snapshot(int, int, double*) const; \
    template psvfwi_status Plan<T>::snapshot_count(int, int*) const; \
    template psvfwi_status Plan<T>::get_records(int, double*) const; \
    template psvfwi_status Plan<T>::get_field(int, int, double*) const; \
    template psvfwi_status Plan<T>::get_store(int, double*) const; \
    template psvfwi_status Plan<T>::get_gradient(int, double*, double*, double*) const; \
    template psvfwi_status Plan<T>::download(const T*, size_t, double*) const;
PSV_FOR_EACH_PRECISION(PSV_INSTANTIATE)

}  // namespace psvplan
