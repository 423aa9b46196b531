// GPUMemoryCtrl.h -- Include/GPUMemoryCtrl.h:6-62 over the C ABI (bool API,
// refuses a second allocation, bounds-checked copies).
#pragma once
#include <cstddef>

#include "lct.h"

lct_ctx *lct_default_ctx();  // the process-wide context of the drop-in classes

template <class T>
class GPUMemoryCtrl
{
private:
    T *pData;
    size_t s;

public:
    GPUMemoryCtrl() : pData(nullptr), s(0) {}
    ~GPUMemoryCtrl() { free_memory(); }
    GPUMemoryCtrl(const GPUMemoryCtrl &) = delete;
    GPUMemoryCtrl &operator=(const GPUMemoryCtrl &) = delete;
    bool allocate_memory(const size_t eleNum = 1)
    {
        if (s != 0) return false;
        void *p = nullptr;
        if (lct_device_malloc(lct_default_ctx(), &p, eleNum * sizeof(T)) != LCT_OK) return false;
        pData = static_cast<T *>(p);
        s = eleNum;
        return true;
    }
    bool free_memory()
    {
        if (s != 0) {
            if (lct_device_free(lct_default_ctx(), pData) != LCT_OK) return false;
            pData = nullptr;
            s = 0;
        }
        return true;
    }
    bool copy_from_CPU(const T *pDataCPU, const size_t eleNum, const size_t offset = 0)
    {
        if (offset > s || s - offset < eleNum) return false;
        return lct_memcpy_h2d(lct_default_ctx(), pData + offset, pDataCPU, eleNum * sizeof(T)) == LCT_OK;
    }
    bool copy_to_CPU(T *pDataCPU, const size_t eleNum, const size_t offset = 0)
    {
        if (offset > s || s - offset < eleNum) return false;
        return lct_memcpy_d2h(lct_default_ctx(), pDataCPU, pData + offset, eleNum * sizeof(T)) == LCT_OK;
    }
    T *get_pointer(const size_t offset = 0) { return pData + offset; }
    size_t size() const { return s; }
};
