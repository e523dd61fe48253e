// CPU check of the thrust_func.h drop-ins (include/thrust_func.h): a device
// call that libsegclip rejects throws std::runtime_error instead of returning
// silently, and the host stable sort orders (key, payload) pairs by key with
// ties in input order.  No GPU call is made: the rejected sizes fail in the
// argument checks.  Prints "ok" and exits 0 on success.
#include "../../include/thrust_func.h"

#include <cstdio>
#include <stdexcept>

template <class F> static bool throws(F f) {
    try {
        f();
    } catch (const std::runtime_error &) {
        return true;
    }
    return false;
}

int main() {
    double dk[1] = {0.0};
    int di[1] = {0}, ik[1] = {0};
    long lk[1] = {0};
    bool ok = true;
    ok &= throws([&] { thrustDeviceStableSortByKey(dk, di, -1); });
    ok &= throws([&] { thrustDeviceStableSortByKey(ik, di, -1); });
    ok &= throws([&] { thrustDeviceStableSortByKey(lk, di, -1); });
    ok &= throws([&] { thrustDeviceExclusiveScan(ik, -1); });
    ok &= throws([&] { thrustDeviceExclusiveScan(lk, -1); });
    if (!ok) {
        std::printf("a rejected device call did not throw\n");
        return 1;
    }
    double k[8] = {3.0, -1.0, 3.0, 0.0, -0.0, 2.5, -1.0, 0.0};
    int v[8] = {0, 1, 2, 3, 4, 5, 6, 7};
    thrustHostStableSortByKey(k, v, 8);
    const double wk[8] = {-1.0, -1.0, 0.0, -0.0, 0.0, 2.5, 3.0, 3.0};
    const int wv[8] = {1, 6, 3, 4, 7, 5, 0, 2};
    for (int i = 0; i < 8; ++i)
        if (k[i] != wk[i] || v[i] != wv[i]) {
            std::printf("host sort: position %d holds (%g, %d)\n", i, k[i], v[i]);
            return 1;
        }
    std::printf("ok\n");
    return 0;
}
