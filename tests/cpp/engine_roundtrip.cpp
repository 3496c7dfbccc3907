// The C++ layer's Threefry4x64 state transfer (include/vsmcx.hpp) against
// vSMC's own engine: the state read from a vsmc::Threefry4x64 and written to
// another engine must reproduce the first engine's stream exactly, at every
// buffer position.  CPU only; needs the reference headers.
#include <cstdio>
#include <sstream>
#include <string>

#include <vsmc/rng/threefry.hpp>

#include "vsmcx.hpp"

int main()
{
    int failures = 0;
    for (unsigned seed : {1u, 12345u, 4242u}) {
        for (int draws = 0; draws < 9; ++draws) {
            vsmc::Threefry4x64 e(seed);
            for (int i = 0; i < draws; ++i) e();
            const vsmcx::Threefry4x64State s = vsmcx::Threefry4x64State::from(e);
            vsmc::Threefry4x64 f(seed + 99);
            s.to(f);
            std::stringstream a, b;
            a << e;
            b << f;
            bool ok = a.str() == b.str();
            for (int i = 0; i < 11 && ok; ++i) ok = e() == f();
            if (!ok) {
                ++failures;
                std::printf("seed %u after %d draws: state transfer differs\n", seed, draws);
            }
        }
    }
    std::printf("engine round trips, failures %d\n", failures);
    return failures ? 1 : 0;
}
