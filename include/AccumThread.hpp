// AccumThread.hpp — the accumulation thread of the reference's asynchronous
// stage loop (AccumThreadHandle, include/hns_spgemm.cuh:32-130, and
// accum_thread_loop, src/hns_spgemm.cu:34-65): while the main thread
// multiplies the tiles of stage s+1, a second host thread adds the product
// of stage s into the partial C.
//
// MI355X mapping: the accumulation thread works on its own library context,
// i.e. its own HIP stream, created by the handle.  start_accum() marks the
// product on the main thread's stream with an event, and the accumulation
// thread's stream waits for that event before its add, so the add of stage
// s runs on the device beside the multiply of stage s+1.  After an add the
// accumulation thread synchronises its stream and leaves the product to be
// released by the main thread (whose context owns it) at its next
// wait_until_idle() or when the handle is destroyed; no context is ever
// used by two threads.
//
// Protocol (unchanged): IDLE / ACTIVE / EMPTY on an atomic flag; the main
// thread calls wait_until_idle() before handing over the next product
// (start_accum) or an empty stage (signal_empty); the accumulation thread
// runs n_accum_total rounds.  The handle, and with it the accumulation
// context whose pool holds a LocalMatrix accumulator, must outlive the
// caller's last use of the accumulated C.
#pragma once

#include <atomic>
#include <optional>
#include <utility>

#include "AccumCSX.hpp"
#include "KokkosWrap.hpp"

template <typename IT, typename VT>
struct AccumThreadHandle {
    using LocalMatrix = KokkosWrap::LocalMatrix<IT, IT, VT>;

    std::optional<LocalMatrix> C_prod;  // handed over, not yet added
    std::optional<LocalMatrix> C_done;  // added, waiting for the main thread to release it
    LocalMatrix *C_local = nullptr;
    AccumCSX<IT, VT> *C_local_csx = nullptr;
    AccumCSX<IT, VT> *C_accum_csx = nullptr;
    KokkosWrap::Handle *handle = nullptr;
    std::atomic<int> *ready_flag;
    int n_accum_total;
    int dev_id;
    bool empty_matrix = true;
    tr_ctx_t accum_ctx = nullptr;       // the accumulation thread's context (stream)
    tr_event_t product_ready = nullptr;  // recorded on the main stream by start_accum

    static constexpr int IDLE = 1;
    static constexpr int ACTIVE = 2;
    static constexpr int EMPTY = 3;

    // double-buffered partial C (hns_spgemm.cuh:53-59)
    AccumThreadHandle(KokkosWrap::Handle *h, AccumCSX<IT, VT> *local_csx, AccumCSX<IT, VT> *accum_csx,
                      int _n_accum_total, int _dev_id)
        : C_local_csx(local_csx), C_accum_csx(accum_csx), handle(h), n_accum_total(_n_accum_total),
          dev_id(_dev_id) {
        init();
    }

    // LocalMatrix accumulator (hns_spgemm.cuh:61-66)
    AccumThreadHandle(LocalMatrix *local, int _n_accum_total, int _dev_id)
        : C_local(local), n_accum_total(_n_accum_total), dev_id(_dev_id) {
        init();
    }

    AccumThreadHandle(const AccumThreadHandle &) = delete;
    AccumThreadHandle &operator=(const AccumThreadHandle &) = delete;

    ~AccumThreadHandle() {
        release_done();
        if (product_ready) (void)tr_event_destroy(product_ready);
        if (accum_ctx) (void)tr_ctx_destroy(accum_ctx);
        delete ready_flag;
    }

    // accumulation thread: C += C_prod_ (hns_spgemm.cuh:68-98)
    void spadd(LocalMatrix &C_prod_) {
        if (handle == nullptr) {
            LocalMatrix::spadd(C_prod_, *C_local);
        } else {
            if (C_prod_.storage.nnz == 0) return;
            accumulate_into(handle, C_prod_.storage, &C_local_csx, &C_accum_csx);
        }
    }

    // main thread
    void wait_until_idle() {
        ready_flag->wait(ACTIVE, std::memory_order_acquire);
        release_done();
    }

    void start_accum(LocalMatrix &&C_prod_) {
        empty_matrix = false;
        KokkosWrap::detail::check(tr_event_record(KokkosWrap::context(), product_ready), "start_accum");
        C_prod.emplace(std::move(C_prod_));
        ready_flag->store(ACTIVE, std::memory_order_release);
        ready_flag->notify_all();
    }

    void signal_empty() {
        empty_matrix = true;
        ready_flag->store(ACTIVE, std::memory_order_release);
        ready_flag->notify_all();
    }

    bool is_empty() const { return empty_matrix; }

    // releases the product of the last finished add (main thread)
    void release_done() {
        if (C_done) {
            C_done->freeBuffers();
            C_done.reset();
        }
    }

  private:
    void init() {
        ready_flag = new std::atomic<int>(IDLE);
        KokkosWrap::detail::check_plain(tr_ctx_create(&accum_ctx, dev_id, nullptr), "AccumThreadHandle: context");
        KokkosWrap::detail::check_plain(tr_event_create(&product_ready), "AccumThreadHandle: event");
    }
};

// src/hns_spgemm.cu:34-65
template <typename IT, typename VT>
void accum_thread_loop(AccumThreadHandle<IT, VT> &handle) {
    KokkosWrap::bind_context(handle.accum_ctx);
    int ndone = 0;
    while (ndone < handle.n_accum_total) {
        handle.ready_flag->wait(AccumThreadHandle<IT, VT>::IDLE, std::memory_order_acquire);
        if (!handle.is_empty()) {
            KokkosWrap::LocalMatrix<IT, IT, VT> C_prod = std::move(*(handle.C_prod));
            handle.C_prod.reset();
            // the product was made on the main thread's stream
            KokkosWrap::detail::check(tr_ctx_wait_event(handle.accum_ctx, handle.product_ready), "accum wait");
            handle.spadd(C_prod);
            KokkosWrap::detail::check(tr_ctx_synchronize(handle.accum_ctx), "accum synchronize");
            handle.C_done.emplace(std::move(C_prod));
        }
        handle.ready_flag->store(AccumThreadHandle<IT, VT>::IDLE, std::memory_order_release);
        handle.ready_flag->notify_all();
        ++ndone;
    }
}
