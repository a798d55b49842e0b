/*
 * api_ctx.h -- libsais contexts of the C wrappers (api.c, api16.c) and the engine
 * context each one owns (engine.hip).  Internal to the library.
 */
#ifndef SAGPU_API_CTX_H
#define SAGPU_API_CTX_H 1

/* A context handle: an engine context (device stream + pool of device buffers)
 * created on the handle's first call. */
__attribute__((visibility("hidden"))) void *sagpu_host_token_new(void);
__attribute__((visibility("hidden"))) void sagpu_host_token_free(void *token);
/* Makes the calling thread's engine calls use the handle's context until the
 * previous binding is restored; returns that previous binding (NULL: none). */
__attribute__((visibility("hidden"))) void *sagpu_host_bind(const void *token);

/* Returns `call` evaluated with the engine bound to context `ctx` (non-NULL). */
#define WITH_CTX(ctx, type, call)                   \
    do {                                            \
        void *prev_ = sagpu_host_bind(ctx);         \
        type rc_ = (call);                          \
        (void)sagpu_host_bind(prev_);               \
        return rc_;                                 \
    } while (0)

#endif
