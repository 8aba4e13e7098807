// Compiles the reference's CPU-only src/test_enum.cpp in place (its path is
// passed as REF_TEST_ENUM) and exposes its three_type() enumeration so the
// oracle's restatement can be compared with the reference's own code.
// Test infrastructure only; the output lives in oracle/_ref/ (not committed).
#define main ref_test_enum_main
#include REF_TEST_ENUM
#undef main

extern "C" long long ref_three_type(int N, int i0, int j0, int* out) {
    std::set<Q4> s = three_type(N, i0, j0);
    long long n = 0;
    for (const Q4& q : s) {
        out[4 * n + 0] = std::get<0>(q);
        out[4 * n + 1] = std::get<1>(q);
        out[4 * n + 2] = std::get<2>(q);
        out[4 * n + 3] = std::get<3>(q);
        n++;
    }
    return n;
}

extern "C" int ref_test_enum_run(void) { return ref_test_enum_main(); }
