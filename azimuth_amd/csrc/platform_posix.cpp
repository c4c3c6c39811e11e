// platform_posix.cpp -- the reference's platform table (platform.h:26-52)
// implemented on POSIX (Linux), in place of platform_win32.cpp.  Plain host
// C++: files through open(2), threads through pthreads, the monotonic clock,
// and the float atomic as a compare-and-swap on the bit pattern.
//
// glibc's <fcntl.h> declares a `struct file_handle` of its own (for
// name_to_handle_at), which clashes in C++ with the reference's typedef of
// that name; this file therefore sees the reference's type under the name
// PlatformFile.  The functions keep C linkage, so callers see file_handle.
#define file_handle PlatformFile
#include "azimuth_platform_posix.h"
#undef file_handle

#include <errno.h>
#include <fcntl.h>
#include <limits.h>
#include <pthread.h>
#include <sched.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <time.h>
#include <unistd.h>

#include <atomic>
#include <cstdint>

namespace {

// One record per thread.  Two owners: the thread itself (until its
// procedure returns) and the handle (until closed); the last one frees it.
// `ended` is set under gThreadLock and announced on gThreadEnded, which all
// waits share, so a wait for any of several threads needs no per-thread
// condition.
struct ThreadRecord {
    pthread_t tid;
    platform_thread_proc *proc;
    void *params;
    bool ended;
    bool closed;
    int owners;
};

pthread_mutex_t gThreadLock = PTHREAD_MUTEX_INITIALIZER;
pthread_cond_t gThreadEnded;
pthread_once_t gCondOnce = PTHREAD_ONCE_INIT;
std::atomic<uint32_t> gNextThreadId{1};

void initCond() {
    pthread_condattr_t a;
    pthread_condattr_init(&a);
    pthread_condattr_setclock(&a, CLOCK_MONOTONIC);
    pthread_cond_init(&gThreadEnded, &a);
    pthread_condattr_destroy(&a);
}

// Drops one owner (caller holds gThreadLock); true when the record is to be freed.
bool release(ThreadRecord *r) { return --r->owners == 0; }

void *trampoline(void *p) {
    ThreadRecord *r = static_cast<ThreadRecord *>(p);
    (void)r->proc(r->params);
    pthread_mutex_lock(&gThreadLock);
    r->ended = true;
    pthread_cond_broadcast(&gThreadEnded);
    const bool last = release(r);
    pthread_mutex_unlock(&gThreadLock);
    if (last) delete r;   // the handle was closed first: the thread was detached
    return nullptr;
}

bool fdValid(PlatformFile h) { return h.Fd >= 0; }

}  // namespace

extern "C" {

b32 PosixGetProcessorCoreCount(u32 *Result) {
    if (!Result) return 0;
    cpu_set_t set;
    CPU_ZERO(&set);
    long n = 0;
    if (sched_getaffinity(0, sizeof(set), &set) == 0) n = CPU_COUNT(&set);
    if (n <= 0) n = sysconf(_SC_NPROCESSORS_ONLN);
    if (n <= 0) return 0;
    *Result = (u32)n;
    return 1;
}

b32 PosixGetAvailableMemory(u64 *Result) {
    if (!Result) return 0;
    FILE *f = fopen("/proc/meminfo", "r");
    if (f) {
        char line[256];
        unsigned long long kb = 0;
        bool found = false;
        while (fgets(line, sizeof(line), f)) {
            if (sscanf(line, "MemAvailable: %llu kB", &kb) == 1) {
                found = true;
                break;
            }
        }
        fclose(f);
        if (found) {
            *Result = (u64)kb * 1024ull;
            return 1;
        }
    }
    const long pages = sysconf(_SC_AVPHYS_PAGES), size = sysconf(_SC_PAGESIZE);
    if (pages <= 0 || size <= 0) return 0;
    *Result = (u64)pages * (u64)size;
    return 1;
}

b32 PosixQueryTimer(timer *Timer) {
    if (!Timer) return 0;
    struct timespec ts;
    if (clock_gettime(CLOCK_MONOTONIC, &ts) != 0) return 0;
    Timer->Counter = (u64)ts.tv_sec * 1000000000ull + (u64)ts.tv_nsec;
    return 1;
}

b32 PosixQueryTimerFrequency(timer_freq *Freq) {
    if (!Freq) return 0;
    Freq->Value = 1000000000ull;
    return 1;
}

b32 PosixMemoryAllocate(void **Ptr, size_t AllocSize) {
    if (!Ptr || AllocSize == 0) return 0;
    *Ptr = malloc(AllocSize);
    return *Ptr != nullptr;
}

b32 PosixMemoryFree(void **Ptr) {
    if (!Ptr || !*Ptr) return 0;
    free(*Ptr);
    *Ptr = nullptr;
    return 1;
}

b32 PosixMemoryCopy(void *Dest, size_t DestSize, void *Src, size_t SrcSize) {
    const size_t n = DestSize < SrcSize ? DestSize : SrcSize;
    if (n) memcpy(Dest, Src, n);
    return n == SrcSize;
}

b32 PosixOpenFile(PlatformFile *Handle, char *Path, u32 FileAccessMode, u32 FileShareMode,
                  file_creation_attrib FileCreationAttrib) {
    (void)FileShareMode;
    if (!Handle || !Path) return 0;
    Handle->Fd = -1;
    const bool rd = (FileAccessMode & FileAccessMode_Read) != 0, wr = (FileAccessMode & FileAccessMode_Write) != 0;
    int flags = rd && wr ? O_RDWR : (wr ? O_WRONLY : O_RDONLY);
    switch (FileCreationAttrib) {
    case FileCreationAttrib_CreateAlways: flags |= O_CREAT | O_TRUNC; break;
    case FileCreationAttrib_CreateIfNotExist: flags |= O_CREAT | O_EXCL; break;
    case FileCreationAttrib_OpenAlways: flags |= O_CREAT; break;
    case FileCreationAttrib_OpenIfExist: break;
    case FileCreationAttrib_OpenAndTruncate: flags |= O_TRUNC; break;
    default: return 0;
    }
    const int fd = open(Path, flags | O_CLOEXEC, 0644);
    if (fd < 0) return 0;
    Handle->Fd = fd;
    return 1;
}

b32 PosixGetFileSize(PlatformFile FileHandle, u64 *FileSize) {
    if (!FileSize || !fdValid(FileHandle)) return 0;
    struct stat st;
    if (fstat(FileHandle.Fd, &st) != 0) return 0;
    *FileSize = (u64)st.st_size;
    return 1;
}

b32 PosixReadFile32(PlatformFile File, void *MemoryPtr, u32 BytesToRead, u32 *BytesRead) {
    if (!fdValid(File) || !MemoryPtr || BytesToRead == 0) return 0;
    u32 done = 0;
    while (done < BytesToRead) {
        const ssize_t r = read(File.Fd, (char *)MemoryPtr + done, BytesToRead - done);
        if (r < 0) {
            if (errno == EINTR) continue;
            if (BytesRead) *BytesRead = done;
            return 0;
        }
        if (r == 0) break;   // end of file
        done += (u32)r;
    }
    if (BytesRead) *BytesRead = done;
    return 1;
}

b32 PosixSetFilePointer(PlatformFile File, s64 SeekDistance, file_seek_pivot Pivot) {
    if (!fdValid(File)) return 0;
    int whence;
    switch (Pivot) {
    case FileSeekPivot_Begin: whence = SEEK_SET; break;
    case FileSeekPivot_Current: whence = SEEK_CUR; break;
    case FileSeekPivot_End: whence = SEEK_END; break;
    default: return 0;
    }
    return lseek(File.Fd, (off_t)SeekDistance, whence) != (off_t)-1;
}

b32 PosixCloseFileHandle(PlatformFile Handle) {
    if (!fdValid(Handle)) return 0;
    return close(Handle.Fd) == 0;
}

b32 PosixCreateThread(thread_handle *Handle, platform_thread_proc *ThreadProc, void *Params, u32 *ThreadID,
                      size_t StackSize) {
    if (!Handle || !ThreadProc) return 0;
    pthread_once(&gCondOnce, initCond);
    Handle->Ptr = nullptr;
    ThreadRecord *r = new ThreadRecord{pthread_t(), ThreadProc, Params, false, false, 2};
    pthread_attr_t attr;
    pthread_attr_init(&attr);
    if (StackSize) pthread_attr_setstacksize(&attr, StackSize < (size_t)PTHREAD_STACK_MIN ? (size_t)PTHREAD_STACK_MIN : StackSize);
    const int e = pthread_create(&r->tid, &attr, trampoline, r);
    pthread_attr_destroy(&attr);
    if (e != 0) {
        delete r;
        return 0;
    }
    Handle->Ptr = r;
    if (ThreadID) *ThreadID = gNextThreadId.fetch_add(1);
    return 1;
}

b32 PosixWaitForThreads(u32 ThreadCount, thread_handle *ThreadHandles, b32 WaitAll, b32 WaitInfinitely,
                        u32 WaitTimeLimit) {
    if (ThreadCount == 0 || !ThreadHandles) return 0;
    for (u32 i = 0; i < ThreadCount; ++i)
        if (!ThreadHandles[i].Ptr) return 0;
    pthread_once(&gCondOnce, initCond);
    struct timespec deadline;
    clock_gettime(CLOCK_MONOTONIC, &deadline);
    deadline.tv_sec += WaitTimeLimit / 1000u;
    deadline.tv_nsec += (long)(WaitTimeLimit % 1000u) * 1000000L;
    if (deadline.tv_nsec >= 1000000000L) {
        deadline.tv_sec += 1;
        deadline.tv_nsec -= 1000000000L;
    }
    pthread_mutex_lock(&gThreadLock);
    for (;;) {
        u32 ended = 0;
        for (u32 i = 0; i < ThreadCount; ++i) ended += static_cast<ThreadRecord *>(ThreadHandles[i].Ptr)->ended ? 1u : 0u;
        if (WaitAll ? ended == ThreadCount : ended > 0) break;
        if (WaitInfinitely) {
            pthread_cond_wait(&gThreadEnded, &gThreadLock);
        } else if (pthread_cond_timedwait(&gThreadEnded, &gThreadLock, &deadline) == ETIMEDOUT) {
            break;   // a timed-out wait is not a failure (WAIT_TIMEOUT != WAIT_FAILED)
        }
    }
    pthread_mutex_unlock(&gThreadLock);
    return 1;
}

b32 PosixCloseThreadHandle(thread_handle Handle) {
    ThreadRecord *r = static_cast<ThreadRecord *>(Handle.Ptr);
    if (!r) return 0;
    pthread_mutex_lock(&gThreadLock);
    if (r->closed) {
        pthread_mutex_unlock(&gThreadLock);
        return 0;
    }
    r->closed = true;
    const bool ended = r->ended;
    const pthread_t tid = r->tid;   // r may be freed by the thread once the lock is dropped
    const bool last = release(r);
    pthread_mutex_unlock(&gThreadLock);
    if (ended) {
        (void)pthread_join(tid, nullptr);   // the procedure has returned: reap it
    } else {
        (void)pthread_detach(tid);          // still running: it frees its record when done
    }
    if (last) delete r;
    return 1;
}

b32 PosixAtomicAddF32(f32 volatile *Dest, f32 Val) {
    if (!Dest) return 0;
    uint32_t *bits = (uint32_t *)Dest;
    uint32_t seen = __atomic_load_n(bits, __ATOMIC_RELAXED);
    for (;;) {
        f32 cur;
        memcpy(&cur, &seen, 4);
        const f32 sum = cur + Val;
        uint32_t want;
        memcpy(&want, &sum, 4);
        if (__atomic_compare_exchange_n(bits, &seen, want, true, __ATOMIC_ACQ_REL, __ATOMIC_RELAXED)) break;
    }
    return 1;
}

void az_platform_posix_api(az_platform_api *Out) {
    if (!Out) return;
    Out->GetProcessorCoreCount = PosixGetProcessorCoreCount;
    Out->GetAvailableMemory = PosixGetAvailableMemory;
    Out->QueryTimer = PosixQueryTimer;
    Out->QueryTimerFrequency = PosixQueryTimerFrequency;
    Out->MemoryAllocate = PosixMemoryAllocate;
    Out->MemoryFree = PosixMemoryFree;
    Out->MemoryCopy = PosixMemoryCopy;
    Out->OpenFile = PosixOpenFile;
    Out->GetFileSize = PosixGetFileSize;
    Out->ReadFile32 = PosixReadFile32;
    Out->SetFilePointer = PosixSetFilePointer;
    Out->CloseFileHandle = PosixCloseFileHandle;
    Out->CreateThread = PosixCreateThread;
    Out->WaitForThreads = PosixWaitForThreads;
    Out->CloseThreadHandle = PosixCloseThreadHandle;
    Out->AtomicAddF32 = PosixAtomicAddF32;
}

}  // extern "C"
