// pm_view.cpp -- interactive host: the reference's GLUT + pixel-buffer viewer
// (simpleGLMain.cpp:64-123 window and frame loop, simplePBO.cpp:35-182 the
// buffer shared with the GPU, callbacksPBO.cpp:47-142 display and keyboard)
// written for X11 + GLX with HIP-GL interop.  The frame goes into a GL pixel
// buffer registered with HIP (hipGraphicsGLRegisterBuffer) and mapped for
// each frame, is copied into a texture and drawn as one quad.
//
// Keys, as in the reference's keyboard handler: Esc quits, a toggles the
// animation, e toggles photon emission (off: the photon maps of the last
// emission are reused while the spheres move), i the interpolated estimate,
// p the participating medium, - and + change the animation step by 0.01, r
// sets it to 0.01.  Added: v toggles the reference's voxel-grid estimate
// (PM_OPT_VOXEL_GRID on the launchers' context).  The window title shows the
// frame rate every 100 frames, as the reference's computeFPS.
//
// --headless runs the same frame loop into a hipMalloc buffer without a
// display: --keys gives one key per frame ('.' for none), and each frame
// prints its state and a checksum of its pixels; --out writes the last frame
// as a binary PPM.
//
//   pm_view [--size WxH] [--headless] [--frames N] [--keys STR] [--out f.ppm]
#define GL_GLEXT_PROTOTYPES 1
#include <GL/gl.h>
#include <GL/glext.h>
#include <GL/glx.h>
#include <X11/Xlib.h>
#include <X11/Xutil.h>
#include <hip/hip_runtime.h>
// after the runtime header, which declares the types it uses
#include <hip/hip_gl_interop.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <chrono>
#include <string>
#include <vector>

#include "../../../include/photon_mapper.h"

namespace {

#define VIEW_HIP(expr)                                                              \
  do {                                                                              \
    hipError_t _e = (expr);                                                         \
    if (_e != hipSuccess) {                                                         \
      fprintf(stderr, "pm_view: %s failed: %s\n", #expr, hipGetErrorString(_e));    \
      exit(EXIT_FAILURE);                                                           \
    }                                                                               \
  } while (0)

// The viewer's switches; initial values as the reference's globals
// (callbacksPBO.cpp:19-28).
struct ViewState {
  bool anim = false;
  bool emit = true;
  bool interp = false;
  bool media = false;
  bool voxel = false;
  float t = 0.0f;
  float inc = 0.1f;
  bool quit = false;
};

// One key press.  Returns true when the frame must be drawn again.
bool apply_key(ViewState& s, unsigned char key) {
  switch (key) {
    case 27: s.quit = true; return false;
    case 'a': case 'A': s.anim = !s.anim; break;
    case 'e': case 'E': s.emit = !s.emit; break;
    case 'i': case 'I': s.interp = !s.interp; break;
    case 'p': case 'P': s.media = !s.media; break;
    case 'v': case 'V': s.voxel = !s.voxel; break;
    case '-': s.inc -= 0.01f; break;
    case '+': s.inc += 0.01f; break;
    case 'r': s.inc = 0.01f; break;
    default: return false;
  }
  return true;
}

// One displayed frame through the reference's three launchers: the random
// numbers once, the emission while emission is on, the render; then the
// animation clock advances.
void run_frame(ViewState& s, pm_uchar4* dptr, unsigned w, unsigned h, bool& first) {
  if (first) {
    launch_init_random_numbers_kernel();
    first = false;
  }
  const int st = pm_set_option(pm_legacy_context(), PM_OPT_VOXEL_GRID, s.voxel ? 1 : 0);
  if (st != PM_OK) {
    fprintf(stderr, "pm_view: pm_set_option failed: %s\n", pm_status_string(st));
    exit(EXIT_FAILURE);
  }
  if (s.emit) launch_emit_photons_kernel(dptr, w, h, s.t, s.interp, s.media);
  launch_photon_mapping_kernel(dptr, w, h, s.t, s.interp, s.media);
  if (s.anim) s.t += s.inc;
}

uint64_t fnv1a(const unsigned char* p, size_t n) {
  uint64_t x = 1469598103934665603ull;
  for (size_t i = 0; i < n; ++i) x = (x ^ p[i]) * 1099511628211ull;
  return x;
}

bool write_ppm(const char* path, const std::vector<pm_uchar4>& px, unsigned w, unsigned h) {
  FILE* fp = fopen(path, "wb");
  if (!fp) return false;
  fprintf(fp, "P6\n%u %u\n255\n", w, h);
  for (const pm_uchar4& q : px) {
    const unsigned char rgb[3] = {q.x, q.y, q.z};
    fwrite(rgb, 1, 3, fp);
  }
  return fclose(fp) == 0;
}

int run_headless(unsigned w, unsigned h, int frames, const std::string& keys, const char* out) {
  pm_uchar4* dptr = nullptr;
  VIEW_HIP(hipMalloc((void**)&dptr, sizeof(pm_uchar4) * w * h));
  std::vector<pm_uchar4> px((size_t)w * h);
  ViewState s;
  bool first = true;
  for (int f = 0; f < frames; ++f) {
    if (f < (int)keys.size() && keys[f] != '.') apply_key(s, (unsigned char)keys[f]);
    if (s.quit) break;
    const float t = s.t;
    run_frame(s, dptr, w, h, first);
    VIEW_HIP(hipMemcpy(px.data(), dptr, sizeof(pm_uchar4) * px.size(), hipMemcpyDeviceToHost));
    printf("frame %d t=%.3f emit=%d anim=%d interp=%d media=%d voxel=%d sum=%016llx\n", f, t, s.emit ? 1 : 0,
           s.anim ? 1 : 0, s.interp ? 1 : 0, s.media ? 1 : 0, s.voxel ? 1 : 0,
           (unsigned long long)fnv1a((const unsigned char*)px.data(), px.size() * sizeof(pm_uchar4)));
    fflush(stdout);
  }
  VIEW_HIP(hipFree(dptr));
  if (out && !write_ppm(out, px, w, h)) {
    fprintf(stderr, "pm_view: cannot write %s\n", out);
    return 1;
  }
  return 0;
}

int run_window(unsigned w, unsigned h) {
  Display* dpy = XOpenDisplay(nullptr);
  if (!dpy) {
    fprintf(stderr, "pm_view: no X display (set DISPLAY, or run with --headless)\n");
    return 1;
  }
  int attr[] = {GLX_RGBA, GLX_DOUBLEBUFFER, GLX_RED_SIZE, 8, GLX_GREEN_SIZE, 8, GLX_BLUE_SIZE, 8, None};
  XVisualInfo* vi = glXChooseVisual(dpy, DefaultScreen(dpy), attr);
  if (!vi) {
    fprintf(stderr, "pm_view: no RGBA double-buffered GLX visual\n");
    XCloseDisplay(dpy);
    return 1;
  }
  const Window root = RootWindow(dpy, vi->screen);
  XSetWindowAttributes swa;
  memset(&swa, 0, sizeof(swa));
  swa.colormap = XCreateColormap(dpy, root, vi->visual, AllocNone);
  swa.event_mask = KeyPressMask | ExposureMask | StructureNotifyMask;
  const Window win = XCreateWindow(dpy, root, 0, 0, w, h, 0, vi->depth, InputOutput, vi->visual,
                                   CWColormap | CWEventMask, &swa);
  XStoreName(dpy, win, "photon mapper");
  Atom wm_delete = XInternAtom(dpy, "WM_DELETE_WINDOW", False);
  XSetWMProtocols(dpy, win, &wm_delete, 1);
  XMapWindow(dpy, win);
  GLXContext ctx = glXCreateContext(dpy, vi, nullptr, True);
  glXMakeCurrent(dpy, win, ctx);

  // the view of simpleGLMain.cpp:100-121: the unit square, no depth test
  glViewport(0, 0, (GLsizei)w, (GLsizei)h);
  glClearColor(0.0f, 0.0f, 0.0f, 1.0f);
  glDisable(GL_DEPTH_TEST);
  glMatrixMode(GL_MODELVIEW);
  glLoadIdentity();
  glMatrixMode(GL_PROJECTION);
  glLoadIdentity();
  glOrtho(0.0, 1.0, 0.0, 1.0, 0.0, 1.0);

  // pixel buffer shared with HIP, and the texture drawn from it
  GLuint pbo = 0, tex = 0;
  glGenBuffers(1, &pbo);
  glBindBuffer(GL_PIXEL_UNPACK_BUFFER, pbo);
  glBufferData(GL_PIXEL_UNPACK_BUFFER, (GLsizeiptr)w * h * 4, nullptr, GL_DYNAMIC_COPY);
  glBindBuffer(GL_PIXEL_UNPACK_BUFFER, 0);
  hipGraphicsResource_t res = nullptr;
  VIEW_HIP(hipGraphicsGLRegisterBuffer(&res, pbo, hipGraphicsRegisterFlagsWriteDiscard));
  glEnable(GL_TEXTURE_2D);
  glGenTextures(1, &tex);
  glBindTexture(GL_TEXTURE_2D, tex);
  glTexImage2D(GL_TEXTURE_2D, 0, GL_RGBA8, (GLsizei)w, (GLsizei)h, 0, GL_RGBA, GL_UNSIGNED_BYTE, nullptr);
  glTexParameteri(GL_TEXTURE_2D, GL_TEXTURE_MIN_FILTER, GL_LINEAR);
  glTexParameteri(GL_TEXTURE_2D, GL_TEXTURE_MAG_FILTER, GL_LINEAR);

  ViewState s;
  bool first = true, redraw = true;
  int fps_count = 0;
  double fps_ms = 0.0;
  while (!s.quit) {
    // events: block for one when nothing is animating
    if (!redraw && !s.anim) {
      XEvent ev;
      XPeekEvent(dpy, &ev);
    }
    while (XPending(dpy)) {
      XEvent ev;
      XNextEvent(dpy, &ev);
      if (ev.type == KeyPress) {
        char buf[8];
        KeySym sym;
        const int n = XLookupString(&ev.xkey, buf, (int)sizeof(buf), &sym, nullptr);
        if (n > 0 && apply_key(s, (unsigned char)buf[0])) redraw = true;
      } else if (ev.type == Expose) {
        redraw = true;
      } else if (ev.type == ClientMessage && (Atom)ev.xclient.data.l[0] == wm_delete) {
        s.quit = true;
      }
    }
    if (s.quit) break;
    if (!redraw && !s.anim) continue;
    const auto t0 = std::chrono::steady_clock::now();
    VIEW_HIP(hipGraphicsMapResources(1, &res, 0));
    void* dptr = nullptr;
    size_t bytes = 0;
    VIEW_HIP(hipGraphicsResourceGetMappedPointer(&dptr, &bytes, res));
    if (bytes < (size_t)w * h * 4) {
      fprintf(stderr, "pm_view: mapped pixel buffer too small (%zu bytes)\n", bytes);
      exit(EXIT_FAILURE);
    }
    run_frame(s, (pm_uchar4*)dptr, w, h, first);
    VIEW_HIP(hipGraphicsUnmapResources(1, &res, 0));
    // texture from the pixel buffer (callbacksPBO.cpp:67-79), then one quad
    // with the first image row at the top (:84-88)
    glBindBuffer(GL_PIXEL_UNPACK_BUFFER, pbo);
    glBindTexture(GL_TEXTURE_2D, tex);
    glTexSubImage2D(GL_TEXTURE_2D, 0, 0, 0, (GLsizei)w, (GLsizei)h, GL_RGBA, GL_UNSIGNED_BYTE, nullptr);
    glBindBuffer(GL_PIXEL_UNPACK_BUFFER, 0);
    glBegin(GL_QUADS);
    glTexCoord2f(0.0f, 1.0f); glVertex3f(0.0f, 0.0f, 0.0f);
    glTexCoord2f(0.0f, 0.0f); glVertex3f(0.0f, 1.0f, 0.0f);
    glTexCoord2f(1.0f, 0.0f); glVertex3f(1.0f, 1.0f, 0.0f);
    glTexCoord2f(1.0f, 1.0f); glVertex3f(1.0f, 0.0f, 0.0f);
    glEnd();
    glXSwapBuffers(dpy, win);
    redraw = false;
    fps_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (++fps_count == 100) {
      char title[96];
      snprintf(title, sizeof(title), "photon mapper: %.1f fps", 1000.0 * fps_count / fps_ms);
      XStoreName(dpy, win, title);
      fps_count = 0;
      fps_ms = 0.0;
    }
  }
  VIEW_HIP(hipGraphicsUnregisterResource(res));
  glDeleteTextures(1, &tex);
  glDeleteBuffers(1, &pbo);
  glXMakeCurrent(dpy, None, nullptr);
  glXDestroyContext(dpy, ctx);
  XDestroyWindow(dpy, win);
  XFree(vi);
  XCloseDisplay(dpy);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  unsigned w = 512, h = 512;  // image_width / image_height of simplePBO.cpp:17-18
  bool headless = false;
  int frames = 1;
  std::string keys;
  const char* out = nullptr;
  for (int i = 1; i < argc; ++i) {
    const bool more = i + 1 < argc;
    if (!strcmp(argv[i], "--headless")) {
      headless = true;
    } else if (!strcmp(argv[i], "--size") && more) {
      if (sscanf(argv[++i], "%ux%u", &w, &h) != 2 || !w || !h) {
        fprintf(stderr, "pm_view: --size wants WxH\n");
        return 2;
      }
    } else if (!strcmp(argv[i], "--frames") && more) {
      frames = atoi(argv[++i]);
    } else if (!strcmp(argv[i], "--keys") && more) {
      keys = argv[++i];
    } else if (!strcmp(argv[i], "--out") && more) {
      out = argv[++i];
    } else {
      fprintf(stderr, "usage: pm_view [--size WxH] [--headless] [--frames N] [--keys STR] [--out f.ppm]\n");
      return 2;
    }
  }
  return headless ? run_headless(w, h, frames, keys, out) : run_window(w, h);
}
