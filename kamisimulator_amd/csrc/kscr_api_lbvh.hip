// kscr_api_lbvh.hip -- the LBVH of the C ABI (kscr_lbvh_*), over the
// builder, refit and query of kscr_lbvh.hip.
#include <cstring>
#include <string>
#include <vector>

#include "kscr_api_internal.h"

using namespace kscr;
using namespace kscr::api;

// ------------------------------------------------------------------ LBVH
struct kscr_lbvh {
    int device = 0;
    LbvhDev d;  // its counter block holds the query's error word at d.count[2]
    int2* pairs = nullptr;
    size_t pair_cap = 0;
    // page-locked copy of the query's counter block (pairs found, redo count,
    // error word): a device-to-host copy into pageable memory goes through a
    // staging buffer and costs tens of microseconds more than the query
    int* h_cnt = nullptr;
    ~kscr_lbvh() {
        lbvh_release(d);
        if (pairs) (void)hipFree(pairs);
        if (h_cnt) (void)hipHostFree(h_cnt);
    }
};

extern "C" {

int kscr_lbvh_create(int device, kscr_lbvh** out) {
    if (!out) return KSCR_ERR_INVALID_ARGUMENT;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return KSCR_ERR_HIP;
    if (device < 0) device = count - 1;
    if (device >= count) return KSCR_ERR_INVALID_ARGUMENT;
    auto* b = new kscr_lbvh;
    b->device = device;
    if (hipSetDevice(device) != hipSuccess) {
        delete b;
        return KSCR_ERR_HIP;
    }
    if (hipHostMalloc((void**)&b->h_cnt, 3 * sizeof(int), hipHostMallocDefault) != hipSuccess) {
        b->h_cnt = nullptr;
        delete b;
        return KSCR_ERR_OUT_OF_MEMORY;
    }
    *out = b;
    return KSCR_OK;
}

void kscr_lbvh_destroy(kscr_lbvh* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    (void)hipDeviceSynchronize();
    delete b;
}

static int lbvh_upload(kscr_lbvh* b, const float* aabbs, int n) {
    if (hipSetDevice(b->device) != hipSuccess) return KSCR_ERR_HIP;
    if (lbvh_reserve(b->d, n) != hipSuccess) return KSCR_ERR_OUT_OF_MEMORY;
    if (hipMemcpy(b->d.aabb, aabbs, (size_t)n * 6 * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
        return KSCR_ERR_HIP;
    b->d.src = nullptr;  // the uploaded copy
    return KSCR_OK;
}

int kscr_lbvh_compute(kscr_lbvh* b, const float* aabbs, int n) {
    if (!b || !aabbs || n <= 0) return KSCR_ERR_INVALID_ARGUMENT;
    int rc = lbvh_upload(b, aabbs, n);
    if (rc) return rc;
    if (lbvh_build(b->d, n, nullptr) != hipSuccess || hipDeviceSynchronize() != hipSuccess) return KSCR_ERR_HIP;
    return KSCR_OK;
}

int kscr_lbvh_refit(kscr_lbvh* b, const float* aabbs) {
    if (!b || !aabbs || !b->d.built) return KSCR_ERR_INVALID_ARGUMENT;
    int rc = lbvh_upload(b, aabbs, b->d.n);
    if (rc) return rc;
    if (lbvh_refit(b->d, nullptr) != hipSuccess || hipDeviceSynchronize() != hipSuccess) return KSCR_ERR_HIP;
    return KSCR_OK;
}

int kscr_lbvh_query(kscr_lbvh* b, int* pairs, size_t cap, size_t* count) {
    if (!b || !count || !b->d.built || (cap > 0 && !pairs)) return KSCR_ERR_INVALID_ARGUMENT;
    if (hipSetDevice(b->device) != hipSuccess) return KSCR_ERR_HIP;
    if (cap > b->pair_cap) {
        if (b->pairs) (void)hipFree(b->pairs);
        b->pairs = nullptr;
        if (hipMalloc((void**)&b->pairs, cap * sizeof(int2)) != hipSuccess) return KSCR_ERR_OUT_OF_MEMORY;
        b->pair_cap = cap;
    }
    if (lbvh_query(b->d, b->pairs, (int)cap, -1, b->d.count + 2, nullptr) != hipSuccess ||
        hipDeviceSynchronize() != hipSuccess)
        return KSCR_ERR_HIP;
    int* cnt = b->h_cnt;  // pairs found, redo count, error word
    if (hipMemcpy(cnt, b->d.count, 3 * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) return KSCR_ERR_HIP;
    const int found = cnt[0], err = cnt[2];
    const size_t m = std::min((size_t)found, cap);
    if (m && hipMemcpy(pairs, b->pairs, m * sizeof(int2), hipMemcpyDeviceToHost) != hipSuccess) return KSCR_ERR_HIP;
    *count = (size_t)found;
    return err ? KSCR_ERR_MAX_COLLISIONS : KSCR_OK;
}

int kscr_lbvh_compute_device(kscr_lbvh* b, const float* d_aabbs, int n, void* stream) {
    if (!b || !d_aabbs || n <= 0) return KSCR_ERR_INVALID_ARGUMENT;
    const hipStream_t s = static_cast<hipStream_t>(stream);
    if (hipSetDevice(b->device) != hipSuccess) return KSCR_ERR_HIP;
    if (lbvh_reserve(b->d, n) != hipSuccess) return KSCR_ERR_OUT_OF_MEMORY;
    b->d.src = d_aabbs;  // read in place until the next compute or refit
    if (lbvh_build(b->d, n, s) != hipSuccess) return KSCR_ERR_HIP;
    return KSCR_OK;
}

int kscr_lbvh_refit_device(kscr_lbvh* b, const float* d_aabbs, void* stream) {
    if (!b || !d_aabbs || !b->d.built) return KSCR_ERR_INVALID_ARGUMENT;
    const hipStream_t s = static_cast<hipStream_t>(stream);
    if (hipSetDevice(b->device) != hipSuccess) return KSCR_ERR_HIP;
    b->d.src = d_aabbs;
    if (lbvh_refit(b->d, s) != hipSuccess) return KSCR_ERR_HIP;
    return KSCR_OK;
}

int kscr_lbvh_query_device(kscr_lbvh* b, int* d_pairs, size_t cap, size_t* count, void* stream) {
    if (!b || !count || !b->d.built || (cap > 0 && !d_pairs) || cap > (size_t)INT32_MAX)
        return KSCR_ERR_INVALID_ARGUMENT;
    const hipStream_t s = static_cast<hipStream_t>(stream);
    if (hipSetDevice(b->device) != hipSuccess) return KSCR_ERR_HIP;
    if (lbvh_query(b->d, reinterpret_cast<int2*>(d_pairs), (int)cap, -1, b->d.count + 2, s) != hipSuccess)
        return KSCR_ERR_HIP;
    int* cnt = b->h_cnt;  // pairs found, redo count, error word: one copy back
    if (hipMemcpyAsync(cnt, b->d.count, 3 * sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return KSCR_ERR_HIP;
    *count = (size_t)cnt[0];
    return cnt[2] ? KSCR_ERR_MAX_COLLISIONS : KSCR_OK;
}

int kscr_lbvh_set_fit_slots(kscr_lbvh* b, int slots) {
    if (!b || slots < 0 || slots > kLbvhFitSlots) return KSCR_ERR_INVALID_ARGUMENT;
    b->d.fit_slots = slots;
    return KSCR_OK;
}

int kscr_lbvh_get(kscr_lbvh* b, uint32_t* codes, uint32_t* ids, uint32_t* links, float* bounds, uint32_t* leaf_parent,
                  float* root, int* max_stack) {
    if (!b || !b->d.built) return KSCR_ERR_INVALID_ARGUMENT;
    const int n = b->d.n;
    if (hipSetDevice(b->device) != hipSuccess) return KSCR_ERR_HIP;
    if (codes && hipMemcpy(codes, b->d.codes, n * 4, hipMemcpyDeviceToHost) != hipSuccess) return KSCR_ERR_HIP;
    if (ids && hipMemcpy(ids, b->d.ids, n * 4, hipMemcpyDeviceToHost) != hipSuccess) return KSCR_ERR_HIP;
    if (leaf_parent && hipMemcpy(leaf_parent, b->d.leaf_parent, n * 4, hipMemcpyDeviceToHost) != hipSuccess)
        return KSCR_ERR_HIP;
    if (root && hipMemcpy(root, b->d.root, 6 * 4, hipMemcpyDeviceToHost) != hipSuccess) return KSCR_ERR_HIP;
    if (max_stack && hipMemcpy(max_stack, b->d.max_depth, 4, hipMemcpyDeviceToHost) != hipSuccess) return KSCR_ERR_HIP;
    if ((links || bounds) && n > 1) {
        std::vector<LbvhNode> nodes(n - 1);
        if (hipMemcpy(nodes.data(), b->d.nodes, (size_t)(n - 1) * sizeof(LbvhNode), hipMemcpyDeviceToHost) != hipSuccess)
            return KSCR_ERR_HIP;
        for (int i = 0; i < n - 1; i++) {
            if (links) {
                links[4 * i] = nodes[i].parent;
                links[4 * i + 1] = nodes[i].left;
                links[4 * i + 2] = nodes[i].right;
                links[4 * i + 3] = nodes[i].fence;
            }
            if (bounds)
                for (int k = 0; k < 12; k++) bounds[12 * i + k] = nodes[i].b[k / 6][k % 6];
        }
    }
    return KSCR_OK;
}

}  // extern "C"

