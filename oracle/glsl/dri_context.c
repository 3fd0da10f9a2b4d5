/*
 * Headless OpenGL 4.5 core context on Mesa's software rasteriser (llvmpipe),
 * created straight through the DRI swrast driver interface — no X server, no
 * EGL, no OSMesa.  Test infrastructure only: it hosts the reference's GLSL
 * compute passes so they can be run on the CPU as the correctness oracle.
 *
 * Exported API (used from Python through ctypes):
 *   int   vo_gl_init(const char *driver_path)   create + bind the context (0 = ok)
 *   void *vo_gl_proc(const char *name)           GL entry point through libglapi
 *   const char *vo_gl_error(void)                last init error
 *
 * The context is bound surfaceless (bindContext(ctx, NULL, NULL)); the passes
 * are compute-only, so no default framebuffer is needed.
 */
#include <dlfcn.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <GL/internal/dri_interface.h>

static char g_err[512];
static void *g_glapi;
static void *(*g_getproc)(const char *);

static void drawable_info(__DRIdrawable *d, int *x, int *y, int *w, int *h, void *lp)
{
    (void)d; (void)lp;
    *x = *y = 0;
    *w = *h = 1;
}
static void put_image(__DRIdrawable *d, int op, int x, int y, int w, int h, char *data, void *lp)
{
    (void)d; (void)op; (void)x; (void)y; (void)w; (void)h; (void)data; (void)lp;
}
static void get_image(__DRIdrawable *d, int x, int y, int w, int h, char *data, void *lp)
{
    (void)d; (void)x; (void)y; (void)lp;
    memset(data, 0, (size_t)w * (size_t)h * 4);
}

static const __DRIswrastLoaderExtension k_loader = {
    {__DRI_SWRAST_LOADER, 1}, drawable_info, put_image, get_image, NULL, NULL, NULL, NULL, NULL};
static const __DRIextension *k_loader_exts[] = {&k_loader.base, NULL};

const char *vo_gl_error(void) { return g_err; }

int vo_gl_init(const char *driver_path)
{
    if (g_getproc) return 0;
    g_glapi = dlopen("libglapi.so.0", RTLD_NOW | RTLD_GLOBAL);
    if (!g_glapi) { snprintf(g_err, sizeof g_err, "libglapi: %s", dlerror()); return -1; }
    void *drv = dlopen(driver_path, RTLD_NOW | RTLD_GLOBAL);
    if (!drv) { snprintf(g_err, sizeof g_err, "driver: %s", dlerror()); return -2; }
    const __DRIextension **(*get_exts)(void) =
        (const __DRIextension **(*)(void))dlsym(drv, "__driDriverGetExtensions_swrast");
    if (!get_exts) { snprintf(g_err, sizeof g_err, "no swrast entry point"); return -3; }
    const __DRIextension **exts = get_exts();
    const __DRIcoreExtension *core = NULL;
    const __DRIswrastExtension *sw = NULL;
    for (int i = 0; exts[i]; i++) {
        if (!strcmp(exts[i]->name, __DRI_CORE)) core = (const __DRIcoreExtension *)exts[i];
        if (!strcmp(exts[i]->name, __DRI_SWRAST)) sw = (const __DRIswrastExtension *)exts[i];
    }
    if (!core || !sw || sw->base.version < 4) {
        snprintf(g_err, sizeof g_err, "driver lacks DRI_Core/DRI_SWRast v4");
        return -4;
    }
    const __DRIconfig **configs = NULL;
    __DRIscreen *scr = sw->createNewScreen2(0, k_loader_exts, exts, &configs, NULL);
    if (!scr || !configs || !configs[0]) { snprintf(g_err, sizeof g_err, "createNewScreen2 failed"); return -5; }
    const uint32_t attribs[] = {__DRI_CTX_ATTRIB_MAJOR_VERSION, 4, __DRI_CTX_ATTRIB_MINOR_VERSION, 5};
    unsigned cerr = 0;
    __DRIcontext *ctx = sw->createContextAttribs(scr, __DRI_API_OPENGL_CORE, configs[0], NULL, 2,
                                                 attribs, &cerr, NULL);
    if (!ctx) { snprintf(g_err, sizeof g_err, "createContextAttribs error %u", cerr); return -6; }
    if (!core->bindContext(ctx, NULL, NULL)) { snprintf(g_err, sizeof g_err, "bindContext failed"); return -7; }
    g_getproc = (void *(*)(const char *))dlsym(g_glapi, "_glapi_get_proc_address");
    if (!g_getproc) { snprintf(g_err, sizeof g_err, "no _glapi_get_proc_address"); return -8; }
    return 0;
}

void *vo_gl_proc(const char *name) { return g_getproc ? g_getproc(name) : NULL; }
