// Drives the host layer of libzlfft (C-ABI argument checks, plan creation and
// its clean-up on failure, ragged-batch validation and item planning, the
// tuning-file parser) in a build with AddressSanitizer and
// UndefinedBehaviorSanitizer on the host side.  Runs without a GPU: calls that
// reach the device fail with ZLFFT_ERR_HIP, which exercises their clean-up.
// Exit status 0 when every check holds; the sanitizers abort on an error.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "zlfft.h"

static int failures = 0;
#define CHECK(cond)                                                  \
    do {                                                             \
        if (!(cond)) {                                               \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #cond, __LINE__); \
            ++failures;                                              \
        }                                                            \
    } while (0)

int main(int argc, char** argv) {
    // a tuning file with malformed, oversized and out-of-range lines
    const char* tune = argc > 1 ? argv[1] : nullptr;
    if (tune) setenv("ZLFFT_TUNE_FILE", tune, 1);

    zlfft_plan p = nullptr;
    CHECK(zlfft_plan_create(nullptr, ZLFFT_KIND_CFFT, ZLFFT_PREC_F32, 4) == ZLFFT_ERR_NULL);
    CHECK(zlfft_plan_create(&p, 7, ZLFFT_PREC_F32, 10) == ZLFFT_ERR_INVALID_ARG && !p);
    CHECK(zlfft_plan_create(&p, ZLFFT_KIND_CFFT, 99, 10) == ZLFFT_ERR_UNSUPPORTED && !p);
    CHECK(zlfft_plan_create(&p, ZLFFT_KIND_RFFT, ZLFFT_PREC_F32, 0) == ZLFFT_ERR_INVALID_ORDER && !p);
    CHECK(zlfft_plan_create(&p, ZLFFT_KIND_CFFT, ZLFFT_PREC_F32, 29) == ZLFFT_ERR_INVALID_ORDER && !p);
    CHECK(zlfft_plan_create(&p, ZLFFT_KIND_CFFT, ZLFFT_PREC_F32, (size_t)-1) == ZLFFT_ERR_INVALID_ORDER && !p);
    CHECK(zlfft_plan_create(&p, ZLFFT_KIND_RFFT, ZLFFT_PREC_F64, 29) == ZLFFT_ERR_INVALID_ORDER && !p);
    CHECK(zlfft_plan_create(&p, ZLFFT_KIND_RFFT, ZLFFT_PREC_F16_MFMA_SPLIT, 12) == ZLFFT_ERR_UNSUPPORTED && !p);
    CHECK(zlfft_plan_create(&p, ZLFFT_KIND_CFFT, ZLFFT_PREC_BF16_MFMA, 10) == ZLFFT_ERR_UNSUPPORTED && !p);
    CHECK(std::strlen(zlfft_last_error()) > 0);

    // valid plans of every path: created (GPU present) or refused at the
    // device (no GPU), with everything built so far released
    const int orders[] = {0, 4, 10, 14, 15, 20, 28};
    for (int order : orders) {
        for (int prec : {ZLFFT_PREC_F32, ZLFFT_PREC_F64}) {
            for (int kind : {ZLFFT_KIND_CFFT, ZLFFT_KIND_RFFT}) {
                if (kind == ZLFFT_KIND_RFFT && order == 0) continue;
                if (prec == ZLFFT_PREC_F64 && order > 27) continue;
                const int rc = zlfft_plan_create(&p, kind, prec, (size_t)order);
                CHECK(rc == ZLFFT_OK || rc == ZLFFT_ERR_HIP);
                if (rc == ZLFFT_OK) {
                    CHECK(zlfft_plan_get_order(p) == (size_t)order);
                    CHECK(zlfft_plan_destroy(p) == ZLFFT_OK);
                } else {
                    CHECK(!p && std::strstr(zlfft_last_error(), "hip") != nullptr);
                }
                p = nullptr;
            }
        }
    }

    // calls on a NULL plan and NULL data
    CHECK(zlfft_plan_destroy(nullptr) == ZLFFT_ERR_NULL);
    CHECK(zlfft_plan_get_size(nullptr) == 0 && zlfft_plan_get_order(nullptr) == 0);
    CHECK(zlfft_cfft_exec(nullptr, 0, 0, nullptr, nullptr, 0, nullptr, nullptr, 1, nullptr) == ZLFFT_ERR_NULL);
    CHECK(zlfft_cfft_exec_noalloc(nullptr, 0, 0, nullptr, nullptr, 0, nullptr, nullptr, 1, nullptr) == ZLFFT_ERR_NULL);
    CHECK(zlfft_rfft_forward_exec(nullptr, nullptr, 0, nullptr, nullptr, 1, nullptr) == ZLFFT_ERR_NULL);
    CHECK(zlfft_rfft_forward_sqr_mag_exec(nullptr, nullptr, nullptr, 1, nullptr) == ZLFFT_ERR_NULL);
    CHECK(zlfft_rfft_backward_exec(nullptr, 0, nullptr, nullptr, nullptr, 1, nullptr) == ZLFFT_ERR_NULL);
    CHECK(zlfft_cfft_execute(nullptr, 0, 0, nullptr, nullptr, 0, nullptr, nullptr) == ZLFFT_ERR_NULL);
    CHECK(zlfft_rfft_forward(nullptr, nullptr, 0, nullptr, nullptr) == ZLFFT_ERR_NULL);
    CHECK(zlfft_rfft_forward_sqr_mag(nullptr, nullptr, nullptr) == ZLFFT_ERR_NULL);
    CHECK(zlfft_rfft_backward(nullptr, 0, nullptr, nullptr, nullptr) == ZLFFT_ERR_NULL);
    CHECK(zlfft_cfft_execute_batch(nullptr, 0, 0, nullptr, nullptr, 0, nullptr, nullptr, 3) == ZLFFT_ERR_NULL);
    CHECK(zlfft_rfft_stft_exec(nullptr, nullptr, 1, nullptr, 1, 0, 0, nullptr, nullptr, 1.f, nullptr) ==
          ZLFFT_ERR_NULL);
    CHECK(zlfft_plan_copy_rfft_twiddles(nullptr, nullptr, 0) == ZLFFT_ERR_NULL);
    CHECK(zlfft_ragged_exec(nullptr, nullptr) == ZLFFT_ERR_NULL);
    CHECK(zlfft_ragged_destroy(nullptr) == ZLFFT_ERR_NULL);

    // ragged batches: every validation branch, then valid sets that are
    // planned on the host (items per order and variant) before the device
    zlfft_ragged r = nullptr;
    static float buf[2 * 32768];
    const zlfft_ragged_desc good = {ZLFFT_KIND_CFFT, 10, ZLFFT_FORWARD, ZLFFT_AOS, buf, nullptr, buf, nullptr};
    CHECK(zlfft_ragged_create(nullptr, &good, 1) == ZLFFT_ERR_NULL);
    CHECK(zlfft_ragged_create(&r, nullptr, 1) == ZLFFT_ERR_NULL && !r);
    auto bad = [&](zlfft_ragged_desc d, int code) {
        std::vector<zlfft_ragged_desc> v(5, good);
        v[3] = d;
        const int rc = zlfft_ragged_create(&r, v.data(), v.size());
        CHECK(rc == code && !r);
        CHECK(std::strstr(zlfft_last_error(), "descriptor 3") != nullptr);
    };
    zlfft_ragged_desc d = good;
    d.layout = 5;
    bad(d, ZLFFT_ERR_INVALID_ARG);
    d = good;
    d.kind = 3;
    bad(d, ZLFFT_ERR_INVALID_ARG);
    d = good;
    d.direction = ZLFFT_FORWARD_SQR_MAG;
    bad(d, ZLFFT_ERR_INVALID_ARG);
    d = good;
    d.kind = ZLFFT_KIND_RFFT;
    d.direction = 9;
    bad(d, ZLFFT_ERR_INVALID_ARG);
    d = good;
    d.order = 2;
    bad(d, ZLFFT_ERR_INVALID_ORDER);
    d = good;
    d.order = 15;
    bad(d, ZLFFT_ERR_INVALID_ORDER);
    d = good;
    d.kind = ZLFFT_KIND_RFFT;
    d.order = 16;
    bad(d, ZLFFT_ERR_INVALID_ORDER);
    d = good;
    d.layout = ZLFFT_SOA;  // SoA CFFT needs in1 and out1
    bad(d, ZLFFT_ERR_NULL);
    d = good;
    d.out0 = nullptr;
    bad(d, ZLFFT_ERR_NULL);
    const int bad_logr[15] = {3, 3, 3, 3, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4};  // order 14 needs 32 points
    CHECK(zlfft_ragged_create_config(&r, &good, 1, bad_logr) == ZLFFT_ERR_INVALID_ARG && !r);
    // a valid mixed set of every kind, direction, layout and order
    std::vector<zlfft_ragged_desc> set;
    for (int order = 3; order <= 15; ++order)
        for (int kind : {ZLFFT_KIND_CFFT, ZLFFT_KIND_RFFT})
            for (int dir : {ZLFFT_FORWARD, ZLFFT_BACKWARD, ZLFFT_FORWARD_SQR_MAG})
                for (int lay : {ZLFFT_AOS, ZLFFT_SOA}) {
                    if (kind == ZLFFT_KIND_CFFT && (dir == ZLFFT_FORWARD_SQR_MAG || order > 14)) continue;
                    if (kind == ZLFFT_KIND_RFFT && order < 4) continue;
                    for (int rep = 0; rep < 3; ++rep)
                        set.push_back({kind, order, dir, lay, buf, buf + 32768, buf, buf + 32768});
                }
    for (int pass = 0; pass < 2; ++pass) {
        const int rc = zlfft_ragged_create(&r, set.data(), set.size());
        CHECK(rc == ZLFFT_OK || rc == ZLFFT_ERR_HIP);
        if (rc == ZLFFT_OK) {
            CHECK(zlfft_ragged_destroy(r) == ZLFFT_OK);
            r = nullptr;
        }
    }
    CHECK(zlfft_ragged_create(&r, set.data(), 0) == ZLFFT_OK || std::strlen(zlfft_last_error()) > 0);
    if (r) zlfft_ragged_destroy(r);

    std::printf("host_abi_driver: %d failed checks\n", failures);
    return failures ? 1 : 0;
}
