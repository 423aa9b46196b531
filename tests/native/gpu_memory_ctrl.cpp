// GPUMemoryCtrl<T> (Include/GPUMemoryCtrl.h:6-62) semantics through the drop-in
// header: a second allocation is refused, copies outside the allocation are
// refused (also when the offset alone is past the end), data round-trips, the
// object can allocate again after free_memory().  Prints "ok" and exits 0 on
// success; prints the failed check and exits 1 otherwise.
#include <cstdio>
#include <vector>

#include "GPUMemoryCtrl.h"

#define CHECK(c)                                                          \
    do {                                                                  \
        if (!(c)) {                                                       \
            std::printf("FAIL line %d: %s\n", __LINE__, #c);              \
            return 1;                                                     \
        }                                                                 \
    } while (0)

int main()
{
    GPUMemoryCtrl<float> m;
    CHECK(m.size() == 0);
    CHECK(m.free_memory());  // nothing to free
    CHECK(m.allocate_memory(1000));
    CHECK(m.size() == 1000);
    CHECK(!m.allocate_memory(10));  // already allocated
    std::vector<float> a(1000), b(1000, -1.0f);
    for (int i = 0; i < 1000; i++) a[i] = 0.5f * float(i) - 7.0f;
    CHECK(m.copy_from_CPU(a.data(), 1000));
    CHECK(!m.copy_from_CPU(a.data(), 1001));
    CHECK(!m.copy_from_CPU(a.data(), 10, 995));
    CHECK(!m.copy_to_CPU(b.data(), 1, 1001));
    CHECK(m.copy_to_CPU(b.data(), 1000));
    for (int i = 0; i < 1000; i++) CHECK(a[i] == b[i]);
    CHECK(m.copy_from_CPU(a.data(), 10, 990));
    std::vector<float> c(10, 0.0f);
    CHECK(m.copy_to_CPU(c.data(), 10, 990));
    for (int i = 0; i < 10; i++) CHECK(c[i] == a[i]);
    CHECK(m.copy_to_CPU(c.data(), 0, 1000));  // an empty copy at the end is in bounds
    CHECK(m.get_pointer(5) == m.get_pointer() + 5);
    CHECK(m.free_memory());
    CHECK(m.size() == 0);
    CHECK(m.allocate_memory(4));
    std::printf("ok\n");
    return 0;
}
