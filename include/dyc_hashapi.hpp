// dyc_hashapi.hpp -- the reference's C++ host class, rebuilt over the C ABI of dyc.h.
//
// Same class name and method names as GPU/dyhash/api.h:9-55 so the reference drivers' call
// sequences compile against it unchanged:
//     hashAPI h(size);  h.hash_insert(k, v, n);  h.hash_search(k, ans, n);
//     h.hash_delete(k, ans, n);  h.resize_up();  h.resize_low();
//     h.set_data_to_gpu(k, v, n);  h.hash_insert_batch();  h.hash_search_batch();
//     h.hash_delete_batch();  h.batch_reset();
// Host arrays in and out, synchronous, like the reference (api.cpp:184-303). Differences:
//   - hash_search writes 0 for a miss exactly as the reference's pre-zeroed output does;
//     hash_search_hits() also returns the hit flags;
//   - hash_delete writes per-key deleted flags into ans (the reference leaves ans unused);
//   - errors throw dyc_error instead of aborting in checkCudaErrors / gtest EXPECT_*.
// Header-only; link with libdyc.so and the HIP runtime (-lamdhip64).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "dyc.h"

#ifndef DYC_HASHAPI_TYPE
#define DYC_HASHAPI_TYPE unsigned int  // dy_hash.h:16
#endif

class dyc_error : public std::runtime_error {
  public:
    dyc_error(int code, const std::string &what) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }

  private:
    int code_;
};

class hashAPI {
  public:
    using TYPE = DYC_HASHAPI_TYPE;
    static constexpr TYPE BATCH_SIZE = 100000;  // test_uilt.h:5
    static constexpr int BATCH_NUM = 1000;      // test_uilt.h:4

    // api.cpp:26-91: `size` is a slot-count hint. `cfg` may carry bounds, kick bound etc.
    explicit hashAPI(TYPE size, const dyc_config *cfg = nullptr)
    {
        check(dyc_create(cfg, size, &t_), "dyc_create");
    }
    ~hashAPI()
    {
        release_pool();
        dyc_destroy(t_);
    }
    hashAPI(const hashAPI &) = delete;
    hashAPI &operator=(const hashAPI &) = delete;

    // Host arrays move through the library's pipelined staging (dyc_*_host): the copy of one
    // chunk overlaps the kernel of the previous one.
    void hash_insert(TYPE *k, TYPE *v, TYPE size)
    {
        check(dyc_insert_host(t_, k, v, size, nullptr), "dyc_insert_host");
    }

    void hash_search(TYPE *k, TYPE *ans, TYPE size) { hash_search_hits(k, ans, nullptr, size); }

    void hash_search_hits(TYPE *k, TYPE *ans, uint8_t *hit, TYPE size)
    {
        check(dyc_find_host(t_, k, ans, hit, size, nullptr), "dyc_find_host");
    }

    void hash_delete(TYPE *k, TYPE *ans, TYPE size)
    {
        std::vector<uint8_t> flags(ans ? size : 0);
        check(dyc_erase_host(t_, k, ans ? flags.data() : nullptr, size, nullptr), "dyc_erase_host");
        if (ans)
            for (TYPE i = 0; i < size; ++i) ans[i] = flags[i];
    }

    // The reference's multi-stream form (GPU/old_version/cuckoo_gtest/api.h:27-31 and
    // gpu_lp_insert(..., stream), cuckoo.cu:865-888): inserts of device arrays enqueued on
    // different streams run at the same time (dyc_insert_async); sync() completes them, and so does
    // any other call.
    void hash_insert_async(TYPE *d_k, TYPE *d_v, TYPE size, hipStream_t stream)
    {
        check(dyc_insert_async(t_, d_k, d_v, size, stream), "dyc_insert_async");
    }
    void sync() { check(dyc_sync(t_), "dyc_sync"); }

    // api.cpp:112-173 ("public for test")
    void resize_up() { check(dyc_resize_up(t_, nullptr), "dyc_resize_up"); }
    void resize_low()
    {
        int rc = dyc_resize_down(t_, nullptr);
        if (rc != DYC_ERR_ODD_BUCKETS) check(rc, "dyc_resize_down");
    }

    // api.cpp:309-344: stage a device pool, then walk it in BATCH_SIZE slices
    bool set_data_to_gpu(TYPE *key, TYPE *value, TYPE size)
    {
        release_pool();
        pool_n_ = size;
        if (hipMalloc(&pool_k_, sizeof(uint32_t) * (size_t)size) != hipSuccess ||
            hipMalloc(&pool_v_, sizeof(uint32_t) * (size_t)size) != hipSuccess ||
            hipMalloc(&check_, sizeof(uint32_t) * (size_t)BATCH_SIZE) != hipSuccess)
            throw dyc_error(DYC_ERR_OOM, "set_data_to_gpu: hipMalloc failed");
        hip(hipMemcpy(pool_k_, key, sizeof(uint32_t) * (size_t)size, hipMemcpyHostToDevice));
        hip(hipMemcpy(pool_v_, value, sizeof(uint32_t) * (size_t)size, hipMemcpyHostToDevice));
        return true;
    }
    // api.cpp:353-369: while the table could not take a whole BATCH_SIZE batch under the upper
    // fill bound, upsize (the reference's api_resize, then the call re-enters itself); then
    // insert the slice. The bound counts BATCH_SIZE whatever the slice holds, and applies
    // whatever the table's auto_resize says.
    void hash_insert_batch()
    {
        TYPE lo, n;
        slice(batch_num_to_op_, lo, n);
        const dyc_config c = config();
        for (dyc_stats s = stats(); (double)(s.count + BATCH_SIZE) > c.upper_bound * (double)s.capacity;
             s = stats())
            check(dyc_resize_up(t_, nullptr), "dyc_resize_up");
        check(dyc_insert(t_, pool_k_ + lo, pool_v_ + lo, n, nullptr), "dyc_insert");
        ++batch_num_to_op_;
    }
    void hash_search_batch()
    {
        TYPE lo, n;
        slice(batch_search_num_to_op_++, lo, n);
        check(dyc_find(t_, pool_k_ + lo, check_, nullptr, n, nullptr), "dyc_find");
        hip(hipDeviceSynchronize());
    }
    // api.cpp:466-506: erase the slice, then downsize once when the count is below the lower
    // fill bound and the capacity above min_capacity (with auto_resize, dyc_erase has applied
    // this same rule already). A subtable with an odd bucket count is left as it is.
    void hash_delete_batch()
    {
        TYPE lo, n;
        slice(batch_del_num_to_op_, lo, n);
        check(dyc_erase(t_, pool_k_ + lo, nullptr, n, nullptr), "dyc_erase");
        ++batch_del_num_to_op_;
        const dyc_config c = config();
        const dyc_stats s = stats();
        if (!c.auto_resize && (double)s.count < c.lower_bound * (double)s.capacity && s.capacity > c.min_capacity) {
            const int rc = dyc_resize_down(t_, nullptr);
            if (rc != DYC_ERR_ODD_BUCKETS) check(rc, "dyc_resize_down");
        }
    }
    void batch_reset() { batch_num_to_op_ = batch_search_num_to_op_ = batch_del_num_to_op_ = 0; }

    dyc_config config() const
    {
        dyc_config c;
        check(dyc_get_config(t_, &c), "dyc_get_config");
        return c;
    }
    dyc_stats stats() const
    {
        dyc_stats s;
        check(dyc_get_stats(t_, &s), "dyc_get_stats");
        return s;
    }
    dyc_table *handle() { return t_; }

  private:
    static void hip(hipError_t e)
    {
        if (e != hipSuccess) throw dyc_error(DYC_ERR_HIP, hipGetErrorString(e));
    }
    static void check(int rc, const char *where)
    {
        if (rc != DYC_OK) throw dyc_error(rc, std::string(where) + ": " + dyc_last_error());
    }
    void slice(int i, TYPE &lo, TYPE &n) const
    {
        if (i >= BATCH_NUM || (uint64_t)i * BATCH_SIZE >= pool_n_)
            throw dyc_error(DYC_ERR_INVALID, "batch index beyond the staged pool");
        lo = (TYPE)(i * BATCH_SIZE);
        n = (pool_n_ - lo < BATCH_SIZE) ? pool_n_ - lo : BATCH_SIZE;
    }
    void release_pool()
    {
        if (pool_k_) (void)hipFree(pool_k_);
        if (pool_v_) (void)hipFree(pool_v_);
        if (check_) (void)hipFree(check_);
        pool_k_ = pool_v_ = check_ = nullptr;
    }

    dyc_table *t_ = nullptr;
    uint32_t *pool_k_ = nullptr, *pool_v_ = nullptr, *check_ = nullptr;
    uint64_t pool_n_ = 0;
    int batch_num_to_op_ = 0, batch_search_num_to_op_ = 0, batch_del_num_to_op_ = 0;
};
