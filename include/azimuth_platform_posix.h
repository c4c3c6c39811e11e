/*
 * azimuth_platform_posix.h -- the reference's platform layer for POSIX
 * (Linux), in place of its Win32 one.
 *
 * The reference reaches the operating system only through 16 function
 * pointers (platform.h:26-52) that SetPlatformAPI(Win32) fills with the
 * Win32Xxx functions of platform_win32.cpp, plus the types of
 * platform_win32.h:3-21.  This header gives the same types and the same 16
 * functions named PosixXxx, so that the reference's own macro
 * SetPlatformAPI(Posix) installs them, and a table az_platform_api for
 * callers that prefer a struct.  Conventions are the reference's: every
 * function returns a b32 success flag and reports through out-parameters;
 * PosixMemoryFree nulls the pointer.
 *
 * The file enums (file_creation_attrib, file_seek_pivot and the access and
 * share bits) belong to the reference's io.h (io.h:9-32); they are declared
 * here too unless AZ_PLATFORM_NO_IO_ENUMS is defined (define it when io.h
 * is included first).
 */
#ifndef AZIMUTH_PLATFORM_POSIX_H
#define AZIMUTH_PLATFORM_POSIX_H

#include <stddef.h>

#include "azimuth_types.h"

#ifdef __cplusplus
extern "C" {
#endif

#ifndef AZ_PLATFORM_NO_IO_ENUMS
typedef enum {
    FileAccessMode_Read  = 0x1,
    FileAccessMode_Write = 0x2
} file_access_mode;
typedef enum {
    FileShareMode_Read   = 0x1,
    FileShareMode_Write  = 0x2,
    FileShareMode_Delete = 0x4
} file_share_mode;
typedef enum {
    FileCreationAttrib_CreateAlways     = 0x01,
    FileCreationAttrib_CreateIfNotExist = 0x02,
    FileCreationAttrib_OpenAlways       = 0x04,
    FileCreationAttrib_OpenIfExist      = 0x08,
    FileCreationAttrib_OpenAndTruncate  = 0x10
} file_creation_attrib;
typedef enum {
    FileSeekPivot_Begin,
    FileSeekPivot_Current,
    FileSeekPivot_End
} file_seek_pivot;
#endif

/* platform_win32.h:3-17 with POSIX contents: a file descriptor (-1: none),
 * an opaque per-thread record, CLOCK_MONOTONIC nanoseconds. */
typedef struct file_handle {
    int Fd;
} file_handle;

typedef struct thread_handle {
    void *Ptr;
} thread_handle;

typedef struct timer {
    u64 Counter;
} timer;

typedef struct timer_freq {
    u64 Value;
} timer_freq;

/* platform_win32.h:19-21: a thread procedure takes its parameter block and
 * returns a 32-bit exit code. */
#define PlatformThreadProc(ProcName) u32 ProcName(void *Parameters)
typedef PlatformThreadProc(platform_thread_proc);

/* Logical processors this process may run on (its affinity mask), as
 * GetNativeSystemInfo's dwNumberOfProcessors counts them on Windows
 * (platform_win32.cpp:257-270). */
b32 PosixGetProcessorCoreCount(u32 *Result);
/* Physical memory available to new allocations (MemAvailable of
 * /proc/meminfo), as ullAvailPhys (platform_win32.cpp:272-289). */
b32 PosixGetAvailableMemory(u64 *Result);
/* Monotonic clock in ticks of PosixQueryTimerFrequency (1e9 per second);
 * platform_win32.cpp:291-325. */
b32 PosixQueryTimer(timer *Timer);
b32 PosixQueryTimerFrequency(timer_freq *Freq);
/* Heap memory (platform_win32.cpp:39-90): allocation fails for a null Ptr
 * or a size of 0; free fails for a null Ptr or *Ptr and nulls *Ptr. */
b32 PosixMemoryAllocate(void **Ptr, size_t AllocSize);
b32 PosixMemoryFree(void **Ptr);
/* Copies min(DestSize, SrcSize) bytes; false when DestSize < SrcSize
 * (platform_win32.cpp:92-104). */
b32 PosixMemoryCopy(void *Dest, size_t DestSize, void *Src, size_t SrcSize);
/* open(2) with the access bits and creation attribute of
 * platform_win32.cpp:106-169 (CreateAlways: create or truncate;
 * CreateIfNotExist: create, fail if it exists; OpenAlways: open or create;
 * OpenIfExist: open, fail if absent; OpenAndTruncate: open and truncate,
 * fail if absent).  Share bits have no POSIX counterpart and are ignored. */
b32 PosixOpenFile(file_handle *Handle, char *Path, u32 FileAccessMode, u32 FileShareMode,
                  file_creation_attrib FileCreationAttrib);
b32 PosixGetFileSize(file_handle FileHandle, u64 *FileSize);
/* Reads up to BytesToRead bytes (fewer at the end of the file); false for
 * an invalid handle, a null buffer or BytesToRead 0 (platform_win32.cpp:189-206). */
b32 PosixReadFile32(file_handle File, void *MemoryPtr, u32 BytesToRead, u32 *BytesRead);
b32 PosixSetFilePointer(file_handle File, s64 SeekDistance, file_seek_pivot Pivot);
b32 PosixCloseFileHandle(file_handle Handle);
/* Starts ThreadProc(Params) on a new thread (StackSize 0: the default);
 * *ThreadID (optional) receives a process-unique id.  The handle stays valid
 * until PosixCloseThreadHandle, whether or not the thread has ended
 * (platform_win32.cpp:327-343). */
b32 PosixCreateThread(thread_handle *Handle, platform_thread_proc *ThreadProc, void *Params, u32 *ThreadID,
                      size_t StackSize);
/* Waits until all (WaitAll) or any of the threads have ended, for at most
 * WaitTimeLimit milliseconds unless WaitInfinitely.  As
 * WaitForMultipleObjects (platform_win32.cpp:345-361), a wait that times out
 * still returns true; false only for invalid arguments. */
b32 PosixWaitForThreads(u32 ThreadCount, thread_handle *ThreadHandles, b32 WaitAll, b32 WaitInfinitely,
                        u32 WaitTimeLimit);
/* Releases the handle; a thread still running keeps running (detached). */
b32 PosixCloseThreadHandle(thread_handle Handle);
/* *Dest += Val as one atomic step: a compare-and-swap loop on the float's
 * bit pattern (platform_win32.cpp:377-395).  The reference compares the
 * old value as a float, which never ends on a NaN and can drop an update
 * between -0 and +0; comparing bit patterns has neither quirk. */
b32 PosixAtomicAddF32(f32 volatile *Dest, f32 Val);

/* The 16 pointers of platform.h:7-23 as one struct. */
typedef struct {
    b32 (*GetProcessorCoreCount)(u32 *);
    b32 (*GetAvailableMemory)(u64 *);
    b32 (*QueryTimer)(timer *);
    b32 (*QueryTimerFrequency)(timer_freq *);
    b32 (*MemoryAllocate)(void **, size_t);
    b32 (*MemoryFree)(void **);
    b32 (*MemoryCopy)(void *, size_t, void *, size_t);
    b32 (*OpenFile)(file_handle *, char *, u32, u32, file_creation_attrib);
    b32 (*GetFileSize)(file_handle, u64 *);
    b32 (*ReadFile32)(file_handle, void *, u32, u32 *);
    b32 (*SetFilePointer)(file_handle, s64, file_seek_pivot);
    b32 (*CloseFileHandle)(file_handle);
    b32 (*CreateThread)(thread_handle *, platform_thread_proc *, void *, u32 *, size_t);
    b32 (*WaitForThreads)(u32, thread_handle *, b32, b32, u32);
    b32 (*CloseThreadHandle)(thread_handle);
    b32 (*AtomicAddF32)(f32 volatile *, f32);
} az_platform_api;

/* Fills *Out with the PosixXxx functions. */
void az_platform_posix_api(az_platform_api *Out);

#ifdef __cplusplus
}
#endif

#endif /* AZIMUTH_PLATFORM_POSIX_H */
