// vSMC's Philox4x32 engine state through text.  Its own operator<< and
// operator>> (include/vsmc/rng/philox.hpp:455-484) disagree on the field
// order; vsmcx::Philox4x32State (include/vsmcx.hpp) reads the written layout,
// round-trips it, and restores an engine that continues the same stream.
// Prints: "<vSMC pair round-trips> <state round-trips> <streams continue equal>"
// as 0/1 flags.
//
// Test infrastructure: compiled by oracle/Makefile (target dropin) against the
// reference headers into oracle/_ref/, run by tests/test_device_header.py.
#include <cstdio>
#include <sstream>

#include <vsmc/rng/philox.hpp>

#include "vsmcx.hpp"

int main()
{
    vsmc::Philox4x32 eng(20250101u);
    eng.discard(6);  // mid-block: two words left of the second block

    std::stringstream ss;
    ss << eng << ' ';
    vsmc::Philox4x32 back;
    ss >> back;
    const int ref_round_trip = back == eng;

    const vsmcx::Philox4x32State s = vsmcx::Philox4x32State::from(eng);
    std::stringstream t;
    s.write(t);
    vsmcx::Philox4x32State s2;
    s2.read(t);
    vsmc::Philox4x32 eng2;
    s2.to(eng2);
    const int state_round_trip = eng2 == eng;

    int same = 1;
    for (int k = 0; k < 20; ++k) same &= eng() == eng2();
    std::printf("%d %d %d\n", ref_round_trip, state_round_trip, same);
    return 0;
}
