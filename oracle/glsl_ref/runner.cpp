// Headless llvmpipe runner for the reference's own GLSL compute shaders.
//
// This program is test infrastructure only. It runs the reference's shader
// files (read as text at run time from <refdir>/shaders, never copied) under
// Mesa llvmpipe through the DRI swrast loader, with no display, and replays
// the dispatch and binding sequence of the reference's step:
//   SPH::Run                     src/SPH.cpp:246-334, including the two UI
//                                highlight passes at :288-296
//   RadixSort::Run / SortBits    src/RadixSort.cpp:124-133, 154-200
//   NeighbourCellFinder::FindNeighbourCells  src/NeighbourCellFinder.cpp:113-148
// It restates the shader-source rule of src/ShaderProgram.cpp:63-93 and the
// generated headers of src/SPH.cpp:28-60, src/RadixSort.cpp:35-39 and
// src/NeighbourCellFinder.cpp:27-30, with h and the grid taken from the job.
//
// Usage: glsl_runner <job.bin> <outdir> [refdir]
// Job layout (little endian): see struct JobHeader below, then pos[N*4] f32,
// vel[N*4] f32, optional end table [gx*gy*gz] i32, optional vorticity [N] f32,
// optional highlight flags [N] u32 (zeros otherwise, as glClearBufferData
// leaves them at src/SPH.cpp:106-109).
// Outputs raw arrays as <outdir>/s<step>_<stage>.bin plus final_*.bin and
// timing.txt (wall clock with glFinish at the five phase boundaries of
// SPH::OutputTiming, src/SPH.cpp:218-240).
#include <GL/glcorearb.h>
#include <GL/internal/dri_interface.h>
#include <dlfcn.h>

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

struct JobHeader {
    char magic[8];  // "PBFJOB2"
    int32_t n, gx, gy, gz;
    float h;
    float params[8];  // one_over_rho_0, epsilon, gravity, timestep, k, scale, xsph c, vorticity eps
    int32_t iterations, vorticity, extforce;
    int32_t nsteps;
    int32_t dump_from, dump_to;  // dump stage outputs for steps in [dump_from, dump_to)
    int32_t has_end_init, has_vort_init;
    int32_t dump_passes;         // also dump every radix pass (bit-exactness of the 2-bit passes)
    int32_t warmup;              // steps excluded from the timing summary
    int32_t has_highlight_init;
};

#define GLFN(T, name) T name = nullptr;
#define GL_FUNCS(X)                                                        \
    X(PFNGLGETSTRINGPROC, glGetString)                                     \
    X(PFNGLGETERRORPROC, glGetError)                                       \
    X(PFNGLFINISHPROC, glFinish)                                           \
    X(PFNGLGENBUFFERSPROC, glGenBuffers)                                   \
    X(PFNGLBINDBUFFERPROC, glBindBuffer)                                   \
    X(PFNGLBUFFERDATAPROC, glBufferData)                                   \
    X(PFNGLBUFFERSUBDATAPROC, glBufferSubData)                             \
    X(PFNGLGETBUFFERSUBDATAPROC, glGetBufferSubData)                       \
    X(PFNGLBINDBUFFERBASEPROC, glBindBufferBase)                           \
    X(PFNGLBINDBUFFERSBASEPROC, glBindBuffersBase)                         \
    X(PFNGLCLEARBUFFERDATAPROC, glClearBufferData)                         \
    X(PFNGLGENTEXTURESPROC, glGenTextures)                                 \
    X(PFNGLBINDTEXTUREPROC, glBindTexture)                                 \
    X(PFNGLTEXBUFFERPROC, glTexBuffer)                                     \
    X(PFNGLTEXIMAGE3DPROC, glTexImage3D)                                   \
    X(PFNGLTEXSUBIMAGE3DPROC, glTexSubImage3D)                             \
    X(PFNGLTEXPARAMETERIPROC, glTexParameteri)                             \
    X(PFNGLTEXPARAMETERIIVPROC, glTexParameterIiv)                         \
    X(PFNGLCLEARTEXIMAGEPROC, glClearTexImage)                             \
    X(PFNGLGETTEXIMAGEPROC, glGetTexImage)                                 \
    X(PFNGLPIXELSTOREIPROC, glPixelStorei)                                 \
    X(PFNGLACTIVETEXTUREPROC, glActiveTexture)                             \
    X(PFNGLBINDIMAGETEXTUREPROC, glBindImageTexture)                       \
    X(PFNGLCREATESHADERPROC, glCreateShader)                               \
    X(PFNGLSHADERSOURCEPROC, glShaderSource)                               \
    X(PFNGLCOMPILESHADERPROC, glCompileShader)                             \
    X(PFNGLGETSHADERIVPROC, glGetShaderiv)                                 \
    X(PFNGLGETSHADERINFOLOGPROC, glGetShaderInfoLog)                       \
    X(PFNGLDELETESHADERPROC, glDeleteShader)                               \
    X(PFNGLCREATEPROGRAMPROC, glCreateProgram)                             \
    X(PFNGLATTACHSHADERPROC, glAttachShader)                               \
    X(PFNGLLINKPROGRAMPROC, glLinkProgram)                                 \
    X(PFNGLGETPROGRAMIVPROC, glGetProgramiv)                               \
    X(PFNGLGETPROGRAMINFOLOGPROC, glGetProgramInfoLog)                     \
    X(PFNGLUSEPROGRAMPROC, glUseProgram)                                   \
    X(PFNGLGETUNIFORMLOCATIONPROC, glGetUniformLocation)                   \
    X(PFNGLPROGRAMUNIFORM1IPROC, glProgramUniform1i)                       \
    X(PFNGLPROGRAMUNIFORM4UIVPROC, glProgramUniform4uiv)                   \
    X(PFNGLDISPATCHCOMPUTEPROC, glDispatchCompute)                         \
    X(PFNGLMEMORYBARRIERPROC, glMemoryBarrier)
GL_FUNCS(GLFN)

void die(const std::string &msg) {
    fprintf(stderr, "glsl_runner: %s\n", msg.c_str());
    exit(2);
}

// ---------------------------------------------------------------- context
void noop_get_drawable_info(__DRIdrawable *, int *x, int *y, int *w, int *h, void *) {
    *x = *y = 0;
    *w = *h = 1;
}
void noop_put_image(__DRIdrawable *, int, int, int, int, int, char *, void *) {}
void noop_get_image(__DRIdrawable *, int, int, int, int, char *, void *) {}

void create_context() {
    if (!dlopen("libglapi.so.0", RTLD_NOW | RTLD_GLOBAL)) die(std::string("dlopen libglapi: ") + dlerror());
    void *drv = dlopen("/usr/lib/x86_64-linux-gnu/dri/swrast_dri.so", RTLD_NOW | RTLD_GLOBAL);
    if (!drv) die(std::string("dlopen swrast_dri: ") + dlerror());
    typedef const __DRIextension **(*GetExt)(void);
    GetExt get = (GetExt)dlsym(drv, "__driDriverGetExtensions_swrast");
    if (!get) die("no __driDriverGetExtensions_swrast");
    const __DRIextension **exts = get();
    const __DRIcoreExtension *core = nullptr;
    const __DRIswrastExtension *swrast = nullptr;
    for (int i = 0; exts[i]; i++) {
        if (!strcmp(exts[i]->name, __DRI_CORE)) core = (const __DRIcoreExtension *)exts[i];
        if (!strcmp(exts[i]->name, __DRI_SWRAST)) swrast = (const __DRIswrastExtension *)exts[i];
    }
    if (!core || !swrast || swrast->base.version < 4) die("swrast driver lacks DRI_Core/DRI_SWRast v4");
    static __DRIswrastLoaderExtension loader;
    memset(&loader, 0, sizeof loader);
    loader.base.name = __DRI_SWRAST_LOADER;
    loader.base.version = 1;
    loader.getDrawableInfo = noop_get_drawable_info;
    loader.putImage = noop_put_image;
    loader.getImage = noop_get_image;
    static const __DRIextension *loader_exts[] = {&loader.base, nullptr};
    const __DRIconfig **configs = nullptr;
    __DRIscreen *screen = swrast->createNewScreen2(0, loader_exts, exts, &configs, nullptr);
    if (!screen || !configs || !configs[0]) die("createNewScreen2 failed");
    uint32_t attribs[] = {__DRI_CTX_ATTRIB_MAJOR_VERSION, 4, __DRI_CTX_ATTRIB_MINOR_VERSION, 3};
    unsigned err = 0;
    __DRIcontext *ctx =
        swrast->createContextAttribs(screen, __DRI_API_OPENGL_CORE, configs[0], nullptr, 2, attribs, &err, nullptr);
    if (!ctx) die("createContextAttribs failed, error " + std::to_string(err));
    if (!core->bindContext(ctx, nullptr, nullptr)) die("bindContext failed");
    typedef void *(*GetProc)(const char *);
    GetProc gp = (GetProc)dlsym(RTLD_DEFAULT, "_glapi_get_proc_address");
    if (!gp) die("no _glapi_get_proc_address");
#define GLLOAD(T, name)                                        \
    name = (T)gp(#name);                                       \
    if (!name) die("missing GL entry point " #name);
    GL_FUNCS(GLLOAD)
}

void check_gl(const char *where) {
    GLenum e = glGetError();
    if (e != GL_NO_ERROR) {
        char b[64];
        snprintf(b, sizeof b, "GL error 0x%x at ", e);
        die(std::string(b) + where);
    }
}

// ---------------------------------------------------------------- programs
std::string g_refdir = "/root/reference";

std::string read_text(const std::string &path) {
    std::ifstream f(path);
    if (!f) die("cannot read " + path);
    std::stringstream s;
    s << f.rdbuf();
    return s.str();
}

// Source rule of ShaderProgram::CompileShader (src/ShaderProgram.cpp:63-93):
// "#version 430 core\n" + header + "\n#line 1\n" + the files concatenated.
GLuint build_program(const std::vector<std::string> &files, const std::string &header) {
    std::string src = "#version 430 core\n" + header + "\n#line 1\n";
    for (const auto &f : files) src += read_text(g_refdir + "/shaders/" + f);
    GLuint sh = glCreateShader(GL_COMPUTE_SHADER);
    const char *p = src.c_str();
    GLint len = (GLint)src.size();
    glShaderSource(sh, 1, &p, &len);
    glCompileShader(sh);
    GLint ok = 0;
    glGetShaderiv(sh, GL_COMPILE_STATUS, &ok);
    if (!ok) {
        char log[8192];
        glGetShaderInfoLog(sh, sizeof log, nullptr, log);
        die("compile " + files.back() + ": " + log);
    }
    GLuint prog = glCreateProgram();
    glAttachShader(prog, sh);
    glLinkProgram(prog);
    glDeleteShader(sh);
    glGetProgramiv(prog, GL_LINK_STATUS, &ok);
    if (!ok) {
        char log[8192];
        glGetProgramInfoLog(prog, sizeof log, nullptr, log);
        die("link " + files.back() + ": " + log);
    }
    return prog;
}

std::string float_literal(float v) {
    char b[64];
    snprintf(b, sizeof b, "%.9g", v);
    std::string s = b;
    if (s.find_first_of(".e") == std::string::npos) s += ".0";
    return s;
}

// ---------------------------------------------------------------- io
void write_file(const std::string &path, const void *data, size_t bytes) {
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) die("cannot write " + path);
    if (bytes && fwrite(data, 1, bytes, f) != bytes) die("short write " + path);
    fclose(f);
}

std::vector<char> read_buffer(GLuint buf, size_t bytes) {
    std::vector<char> v(bytes);
    glBindBuffer(GL_COPY_READ_BUFFER, buf);
    glGetBufferSubData(GL_COPY_READ_BUFFER, 0, bytes, v.data());
    return v;
}

std::vector<char> read_tex3d(GLuint tex, size_t cells) {
    std::vector<char> v(cells * 4);
    glBindTexture(GL_TEXTURE_3D, tex);
    glPixelStorei(GL_PACK_ALIGNMENT, 1);
    glGetTexImage(GL_TEXTURE_3D, 0, GL_RED_INTEGER, GL_INT, v.data());
    return v;
}

uint32_t count_sortbits(uint64_t v) {  // src/RadixSort.cpp:24-30
    uint32_t r = 1;
    while (v >>= 1) r++;
    return r;
}

uint32_t intpow(uint32_t x, uint32_t y) {  // src/RadixSort.cpp:141-152
    uint32_t r = 1;
    while (y) {
        if (y & 1) r *= x;
        y >>= 1;
        x *= x;
    }
    return r;
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 3) die("usage: glsl_runner <job.bin> <outdir> [refdir]");
    if (argc > 3) g_refdir = argv[3];
    const std::string outdir = argv[2];

    FILE *jf = fopen(argv[1], "rb");
    if (!jf) die("cannot open job");
    JobHeader job;
    if (fread(&job, sizeof job, 1, jf) != 1 || strncmp(job.magic, "PBFJOB2", 8) != 0) die("bad job header");
    const int N = job.n;
    const int gx = job.gx, gy = job.gy, gz = job.gz;
    const size_t C = (size_t)gx * gy * gz;
    if (N <= 0 || N % 512) die("N must be a positive multiple of 512 (src/Simulation.cpp:202)");
    std::vector<float> pos(4 * (size_t)N), vel(4 * (size_t)N);
    std::vector<int32_t> end_init;
    std::vector<float> vort_init;
    std::vector<uint32_t> highlight_init((size_t)N, 0u);
    if (fread(pos.data(), 4, pos.size(), jf) != pos.size()) die("short pos");
    if (fread(vel.data(), 4, vel.size(), jf) != vel.size()) die("short vel");
    if (job.has_end_init) {
        end_init.resize(C);
        if (fread(end_init.data(), 4, C, jf) != C) die("short end table");
    }
    if (job.has_vort_init) {
        vort_init.resize(N);
        if (fread(vort_init.data(), 4, N, jf) != (size_t)N) die("short vorticity");
    }
    if (job.has_highlight_init && fread(highlight_init.data(), 4, N, jf) != (size_t)N) die("short highlight");
    fclose(jf);

    create_context();
    fprintf(stderr, "renderer: %s | %s\n", glGetString(GL_RENDERER), glGetString(GL_VERSION));

    // -------- headers (src/SPH.cpp:28-60, src/RadixSort.cpp:35-39, src/NeighbourCellFinder.cpp:27-30)
    std::stringstream grid;
    grid << "const vec3 GRID_SIZE = vec3 (" << gx << ", " << gy << ", " << gz << ");\n"
         << "const ivec3 GRID_HASHWEIGHTS = ivec3 (1, " << gx * gz << ", " << gx << ");\n";
    std::stringstream sph_hdr;
    sph_hdr << grid.str() << "\n"
            << "layout (binding = 2, std140) uniform SPHParameters\n{\n"
            << "  float one_over_rho_0;\n  float epsilon;\n  float gravity;\n  float timestep;\n"
            << "  float tensile_instability_k;\n  float tensile_instability_scale;\n"
            << "  float xsph_viscosity_c;\n  float vorticity_epsilon;\n};\n"
            << "const float h = " << float_literal(job.h) << ";\n\n#define BLOCKSIZE 256\n";
    const uint32_t sort_block = 512, numblocks = (uint32_t)N >> 9;
    std::stringstream sort_hdr;
    sort_hdr << grid.str() << "#define BLOCKSIZE " << sort_block << "\n#define HALFBLOCKSIZE " << sort_block / 2 << "\n";
    std::stringstream ncf_hdr;
    ncf_hdr << grid.str() << "#define BLOCKSIZE 256\n";

    const std::string fe = "sph/foreachneighbour.glsl";
    GLuint p_predict = build_program({fe, "sph/predictpos.glsl"}, sph_hdr.str());
    GLuint p_lambda = build_program({fe, "sph/calclambda.glsl"}, sph_hdr.str());
    GLuint p_updpos = build_program({fe, "sph/updatepos.glsl"}, sph_hdr.str());
    GLuint p_vort = build_program({fe, "sph/vorticity.glsl"}, sph_hdr.str());
    GLuint p_update = build_program({fe, "sph/update.glsl"}, sph_hdr.str());
    GLuint p_highlight = build_program({fe, "sph/highlight.glsl"}, sph_hdr.str());
    GLuint p_clearhighlight = build_program({fe, "sph/clearhighlight.glsl"}, sph_hdr.str());
    GLuint p_counting = build_program({"radixsort/counting.glsl"}, sort_hdr.str());
    GLuint p_blockscan = build_program({"radixsort/blockscan.glsl"}, sort_hdr.str());
    GLuint p_globalsort = build_program({"radixsort/globalsort.glsl"}, sort_hdr.str());
    GLuint p_addblocksum = build_program({"radixsort/addblocksum.glsl"}, sort_hdr.str());
    GLuint p_findcells = build_program({"neighbourcellfinder/findcells.glsl"}, ncf_hdr.str());
    GLuint p_nbrcells = build_program({"neighbourcellfinder/neighbourcells.glsl"}, ncf_hdr.str());
    check_gl("programs");

    // -------- RadixSort objects (src/RadixSort.cpp:32-109)
    const uint32_t numbits = count_sortbits((uint64_t)gx * gy * gz - 1);
    GLuint rs_buf[3];
    glGenBuffers(3, rs_buf);
    GLuint rs_buffer = rs_buf[0], rs_prefix = rs_buf[1], rs_result = rs_buf[2];
    glBindBuffer(GL_SHADER_STORAGE_BUFFER, rs_buffer);
    glBufferData(GL_SHADER_STORAGE_BUFFER, 16 * (size_t)sort_block * numblocks, nullptr, GL_DYNAMIC_COPY);
    glBindBuffer(GL_SHADER_STORAGE_BUFFER, rs_prefix);
    glBufferData(GL_SHADER_STORAGE_BUFFER, 4 * (size_t)sort_block * numblocks, nullptr, GL_DYNAMIC_COPY);
    std::vector<GLuint> blocksums;
    {
        uint32_t numblocksums = 4 * numblocks;
        int levels = (int)ceil(log((double)(((numblocksums + sort_block - 1) / sort_block) * sort_block)) / log((double)sort_block));
        levels++;
        blocksums.resize(levels);
        glGenBuffers(levels, blocksums.data());
        for (int i = 0; i < levels; i++) {
            glBindBuffer(GL_SHADER_STORAGE_BUFFER, blocksums[i]);
            numblocksums = ((numblocksums + sort_block - 1) / sort_block) * sort_block;
            if (numblocksums < 1) numblocksums = 1;
            glBufferData(GL_SHADER_STORAGE_BUFFER, 4 * (size_t)numblocksums, nullptr, GL_DYNAMIC_COPY);
            numblocksums /= sort_block;
        }
    }
    glBindBuffer(GL_SHADER_STORAGE_BUFFER, rs_result);
    glBufferData(GL_SHADER_STORAGE_BUFFER, 16 * (size_t)sort_block * numblocks, nullptr, GL_DYNAMIC_COPY);
    {
        GLuint offs[4] = {0, numblocks, numblocks * 2, numblocks * 3};
        glProgramUniform4uiv(p_counting, glGetUniformLocation(p_counting, "blocksumoffsets"), 1, offs);
        glProgramUniform4uiv(p_globalsort, glGetUniformLocation(p_globalsort, "blocksumoffsets"), 1, offs);
    }
    const GLint counting_bitshift = glGetUniformLocation(p_counting, "bitshift");
    const GLint globalsort_bitshift = glGetUniformLocation(p_globalsort, "bitshift");
    for (GLuint b : blocksums) {
        glBindBuffer(GL_SHADER_STORAGE_BUFFER, b);
        glClearBufferData(GL_SHADER_STORAGE_BUFFER, GL_RGBA32UI, GL_RGBA_INTEGER, GL_UNSIGNED_INT, nullptr);
    }
    check_gl("radixsort setup");

    // -------- NeighbourCellFinder objects (src/NeighbourCellFinder.cpp:24-100)
    GLuint tex[6];
    glGenTextures(6, tex);
    GLuint t_start = tex[0], t_end = tex[1], t_nbr = tex[2], t_lambda = tex[3], t_pos = tex[4], t_vel = tex[5];
    glBindTexture(GL_TEXTURE_3D, t_start);
    glTexImage3D(GL_TEXTURE_3D, 0, GL_R32I, gx, gy, gz, 0, GL_RED_INTEGER, GL_INT, nullptr);
    glTexParameteri(GL_TEXTURE_3D, GL_TEXTURE_MAG_FILTER, GL_NEAREST);
    glTexParameteri(GL_TEXTURE_3D, GL_TEXTURE_MIN_FILTER, GL_NEAREST);
    glTexParameteri(GL_TEXTURE_3D, GL_TEXTURE_WRAP_S, GL_CLAMP_TO_BORDER);
    glTexParameteri(GL_TEXTURE_3D, GL_TEXTURE_WRAP_T, GL_CLAMP_TO_BORDER);
    glTexParameteri(GL_TEXTURE_3D, GL_TEXTURE_WRAP_R, GL_CLAMP_TO_BORDER);
    {
        GLint border[] = {-1, -1, -1, -1};
        glTexParameterIiv(GL_TEXTURE_3D, GL_TEXTURE_BORDER_COLOR, border);
        GLint v = -1;
        glClearTexImage(t_start, 0, GL_RED_INTEGER, GL_INT, &v);
    }
    glBindTexture(GL_TEXTURE_3D, t_end);
    glTexImage3D(GL_TEXTURE_3D, 0, GL_R32I, gx, gy, gz, 0, GL_RED_INTEGER, GL_INT,
                 job.has_end_init ? end_init.data() : nullptr);
    glTexParameteri(GL_TEXTURE_3D, GL_TEXTURE_MAG_FILTER, GL_NEAREST);
    glTexParameteri(GL_TEXTURE_3D, GL_TEXTURE_MIN_FILTER, GL_NEAREST);
    glTexParameteri(GL_TEXTURE_3D, GL_TEXTURE_WRAP_S, GL_CLAMP_TO_EDGE);
    glTexParameteri(GL_TEXTURE_3D, GL_TEXTURE_WRAP_T, GL_CLAMP_TO_EDGE);
    glTexParameteri(GL_TEXTURE_3D, GL_TEXTURE_WRAP_R, GL_CLAMP_TO_EDGE);
    GLuint ncf_buf;
    glGenBuffers(1, &ncf_buf);
    glBindBuffer(GL_SHADER_STORAGE_BUFFER, ncf_buf);
    glBufferData(GL_SHADER_STORAGE_BUFFER, 4 * 12 * (size_t)N, nullptr, GL_DYNAMIC_COPY);
    glBindTexture(GL_TEXTURE_BUFFER, t_nbr);
    glTexBuffer(GL_TEXTURE_BUFFER, GL_RGBA32I, ncf_buf);
    check_gl("ncf setup");

    // -------- SPH objects (src/SPH.cpp:95-150)
    GLuint sb[6];
    glGenBuffers(6, sb);
    GLuint b_lambda = sb[0], b_vort = sb[1], b_pos = sb[2], b_vel = sb[3], b_params = sb[4], b_highlight = sb[5];
    GLuint t_highlight;
    glGenTextures(1, &t_highlight);
    glBindBuffer(GL_SHADER_STORAGE_BUFFER, b_highlight);
    glBufferData(GL_SHADER_STORAGE_BUFFER, 4 * (size_t)N, highlight_init.data(), GL_DYNAMIC_COPY);
    glBindTexture(GL_TEXTURE_BUFFER, t_highlight);
    glTexBuffer(GL_TEXTURE_BUFFER, GL_R32UI, b_highlight);
    glBindBuffer(GL_SHADER_STORAGE_BUFFER, b_lambda);
    glBufferData(GL_SHADER_STORAGE_BUFFER, 4 * (size_t)N, nullptr, GL_DYNAMIC_COPY);
    glBindTexture(GL_TEXTURE_BUFFER, t_lambda);
    glTexBuffer(GL_TEXTURE_BUFFER, GL_R32F, b_lambda);
    glBindBuffer(GL_SHADER_STORAGE_BUFFER, b_vort);
    glBufferData(GL_SHADER_STORAGE_BUFFER, 4 * (size_t)N, job.has_vort_init ? vort_init.data() : nullptr,
                 GL_DYNAMIC_COPY);
    glBindBuffer(GL_SHADER_STORAGE_BUFFER, b_pos);
    glBufferData(GL_SHADER_STORAGE_BUFFER, 16 * (size_t)N, pos.data(), GL_DYNAMIC_COPY);
    glBindTexture(GL_TEXTURE_BUFFER, t_pos);
    glTexBuffer(GL_TEXTURE_BUFFER, GL_RGBA32F, b_pos);
    glBindBuffer(GL_SHADER_STORAGE_BUFFER, b_vel);
    glBufferData(GL_SHADER_STORAGE_BUFFER, 16 * (size_t)N, vel.data(), GL_DYNAMIC_COPY);
    glBindTexture(GL_TEXTURE_BUFFER, t_vel);
    glTexBuffer(GL_TEXTURE_BUFFER, GL_RGBA32F, b_vel);
    glBindBuffer(GL_UNIFORM_BUFFER, b_params);
    glBufferData(GL_UNIFORM_BUFFER, sizeof job.params, job.params, GL_STATIC_DRAW);
    glBindBufferBase(GL_UNIFORM_BUFFER, 2, b_params);
    glProgramUniform1i(p_predict, glGetUniformLocation(p_predict, "extforce"), job.extforce ? 1 : 0);
    check_gl("sph setup");

    // initial contents of the buffers the reference never clears
    write_file(outdir + "/init_vort.bin", read_buffer(b_vort, 4 * (size_t)N).data(), 4 * (size_t)N);
    write_file(outdir + "/init_end.bin", read_tex3d(t_end, C).data(), 4 * C);

    const GLuint groups = (GLuint)N >> 8;
    const int passes = (int)(numbits + 1) >> 1;
    FILE *tf = fopen((outdir + "/timing.txt").c_str(), "w");
    if (!tf) die("cannot write timing.txt");
    fprintf(tf, "# step predict_s sort_s cells_s solver_s update_vorticity_s total_s\n");
    double acc[6] = {0, 0, 0, 0, 0, 0};
    int timed = 0;

    for (int step = 0; step < job.nsteps; step++) {
        const bool dump = step >= job.dump_from && step < job.dump_to;
        auto dumpbuf = [&](const char *stage, GLuint buf, size_t bytes) {
            if (!dump) return;
            char name[256];
            snprintf(name, sizeof name, "%s/s%d_%s.bin", outdir.c_str(), step, stage);
            write_file(name, read_buffer(buf, bytes).data(), bytes);
        };
        glFinish();
        double t0 = now_s();

        // 1. predict (src/SPH.cpp:247-261)
        glBindBufferBase(GL_SHADER_STORAGE_BUFFER, 1, rs_buffer);
        glActiveTexture(GL_TEXTURE0);
        glBindTexture(GL_TEXTURE_BUFFER, t_pos);
        glActiveTexture(GL_TEXTURE1);
        glBindTexture(GL_TEXTURE_BUFFER, t_vel);
        glActiveTexture(GL_TEXTURE0);
        glUseProgram(p_predict);
        glDispatchCompute(groups, 1, 1);
        glMemoryBarrier(GL_SHADER_STORAGE_BARRIER_BIT);
        glFinish();
        double t1 = now_s();
        dumpbuf("predict", rs_buffer, 16 * (size_t)N);

        // 2. RadixSort::Run (src/RadixSort.cpp:124-133, SortBits :154-200)
        for (int i = 0; i < passes; i++) {
            const int bits = 2 * i;
            glProgramUniform1i(p_counting, counting_bitshift, bits);
            glProgramUniform1i(p_globalsort, globalsort_bitshift, bits);
            {
                GLuint bufs[4] = {rs_buffer, rs_prefix, blocksums.front(), rs_result};
                glBindBuffersBase(GL_SHADER_STORAGE_BUFFER, 0, 4, bufs);
            }
            glUseProgram(p_counting);
            glDispatchCompute(numblocks, 1, 1);
            glMemoryBarrier(GL_SHADER_STORAGE_BARRIER_BIT);
            glUseProgram(p_blockscan);
            uint32_t nbs = (4 * numblocks) / sort_block;
            for (size_t l = 0; l + 1 < blocksums.size(); l++) {
                glBindBuffersBase(GL_SHADER_STORAGE_BUFFER, 0, 2, &blocksums[l]);
                glDispatchCompute(nbs > 0 ? nbs : 1, 1, 1);
                nbs /= sort_block;
                glMemoryBarrier(GL_SHADER_STORAGE_BARRIER_BIT);
            }
            glUseProgram(p_addblocksum);
            for (int l = (int)blocksums.size() - 3; l >= 0; l--) {
                uint32_t n2 = (4 * numblocks) / intpow(sort_block, l + 1);
                glBindBuffersBase(GL_SHADER_STORAGE_BUFFER, 0, 2, &blocksums[l]);
                glDispatchCompute(n2 > 0 ? n2 : 1, 1, 1);
                glMemoryBarrier(GL_SHADER_STORAGE_BARRIER_BIT);
            }
            {
                GLuint bufs[2] = {rs_buffer, rs_prefix};
                glBindBuffersBase(GL_SHADER_STORAGE_BUFFER, 0, 2, bufs);
            }
            glUseProgram(p_globalsort);
            glDispatchCompute(numblocks, 1, 1);
            glMemoryBarrier(GL_SHADER_STORAGE_BARRIER_BIT);
            std::swap(rs_result, rs_buffer);
            if (job.dump_passes) {
                char stage[32];
                snprintf(stage, sizeof stage, "pass%d", i);
                dumpbuf(stage, rs_buffer, 16 * (size_t)N);
            }
        }
        glFinish();
        double t2 = now_s();
        dumpbuf("sorted", rs_buffer, 16 * (size_t)N);

        // 3. FindNeighbourCells (src/NeighbourCellFinder.cpp:113-148)
        {
            GLint v = -1;
            glClearTexImage(t_start, 0, GL_RED_INTEGER, GL_INT, &v);
        }
        glBindBufferBase(GL_SHADER_STORAGE_BUFFER, 0, rs_buffer);
        glBindImageTexture(0, t_start, 0, GL_TRUE, 0, GL_WRITE_ONLY, GL_R32I);
        glBindImageTexture(1, t_end, 0, GL_TRUE, 0, GL_WRITE_ONLY, GL_R32I);
        glUseProgram(p_findcells);
        glDispatchCompute(groups, 1, 1);
        glMemoryBarrier(GL_SHADER_STORAGE_BARRIER_BIT | GL_SHADER_IMAGE_ACCESS_BARRIER_BIT);
        glActiveTexture(GL_TEXTURE0);
        glBindTexture(GL_TEXTURE_3D, t_start);
        glActiveTexture(GL_TEXTURE1);
        glBindTexture(GL_TEXTURE_3D, t_end);
        glActiveTexture(GL_TEXTURE0);
        glBindImageTexture(0, t_nbr, 0, GL_FALSE, 0, GL_WRITE_ONLY, GL_RGBA32I);
        glUseProgram(p_nbrcells);
        glDispatchCompute(groups, 1, 1);
        glMemoryBarrier(GL_SHADER_IMAGE_ACCESS_BARRIER_BIT);
        glFinish();
        double t3 = now_s();
        if (dump) {
            char name[256];
            snprintf(name, sizeof name, "%s/s%d_start.bin", outdir.c_str(), step);
            write_file(name, read_tex3d(t_start, C).data(), 4 * C);
            snprintf(name, sizeof name, "%s/s%d_end.bin", outdir.c_str(), step);
            write_file(name, read_tex3d(t_end, C).data(), 4 * C);
            glActiveTexture(GL_TEXTURE0);
            glBindTexture(GL_TEXTURE_3D, t_start);
            glActiveTexture(GL_TEXTURE1);
            glBindTexture(GL_TEXTURE_3D, t_end);
            glActiveTexture(GL_TEXTURE0);
        }
        dumpbuf("nbr", ncf_buf, 4 * 12 * (size_t)N);

        // 4. solver (src/SPH.cpp:277-313)
        glBindBufferBase(GL_SHADER_STORAGE_BUFFER, 1, rs_buffer);
        glActiveTexture(GL_TEXTURE2);
        glBindTexture(GL_TEXTURE_BUFFER, t_nbr);
        glActiveTexture(GL_TEXTURE3);
        glBindTexture(GL_TEXTURE_BUFFER, t_lambda);
        glActiveTexture(GL_TEXTURE0);
        // particle highlighting (src/SPH.cpp:288-296)
        glBindImageTexture(0, t_highlight, 0, GL_FALSE, 0, GL_READ_WRITE, GL_R32UI);
        glUseProgram(p_clearhighlight);
        glDispatchCompute(groups, 1, 1);
        glMemoryBarrier(GL_SHADER_IMAGE_ACCESS_BARRIER_BIT);
        glUseProgram(p_highlight);
        glDispatchCompute(groups, 1, 1);
        glMemoryBarrier(GL_SHADER_IMAGE_ACCESS_BARRIER_BIT);
        dumpbuf("highlight", b_highlight, 4 * (size_t)N);
        glBindImageTexture(0, t_lambda, 0, GL_FALSE, 0, GL_WRITE_ONLY, GL_R32F);
        for (int it = 0; it < job.iterations; it++) {
            glUseProgram(p_lambda);
            glDispatchCompute(groups, 1, 1);
            glMemoryBarrier(GL_SHADER_STORAGE_BARRIER_BIT | GL_TEXTURE_FETCH_BARRIER_BIT |
                            GL_SHADER_IMAGE_ACCESS_BARRIER_BIT);
            char stage[32];
            snprintf(stage, sizeof stage, "lambda%d", it);
            dumpbuf(stage, b_lambda, 4 * (size_t)N);
            glUseProgram(p_updpos);
            glDispatchCompute(groups, 1, 1);
            glMemoryBarrier(GL_SHADER_STORAGE_BARRIER_BIT);
            snprintf(stage, sizeof stage, "pos%d", it);
            dumpbuf(stage, rs_buffer, 16 * (size_t)N);
        }
        glFinish();
        double t4 = now_s();

        // 5. update + vorticity (src/SPH.cpp:315-333)
        glUseProgram(p_update);
        glBindImageTexture(0, t_pos, 0, GL_FALSE, 0, GL_READ_WRITE, GL_RGBA32F);
        glBindImageTexture(1, t_vel, 0, GL_FALSE, 0, GL_READ_WRITE, GL_RGBA32F);
        glDispatchCompute(groups, 1, 1);
        glMemoryBarrier(GL_SHADER_STORAGE_BARRIER_BIT | GL_TEXTURE_FETCH_BARRIER_BIT |
                        GL_SHADER_IMAGE_ACCESS_BARRIER_BIT);
        dumpbuf("upd_pos", b_pos, 16 * (size_t)N);
        dumpbuf("upd_vel", b_vel, 16 * (size_t)N);
        if (job.vorticity) {
            glBindBufferBase(GL_SHADER_STORAGE_BUFFER, 3, b_vort);
            glUseProgram(p_vort);
            glDispatchCompute(groups, 1, 1);
            glMemoryBarrier(GL_SHADER_STORAGE_BARRIER_BIT);
            dumpbuf("vort", b_vort, 4 * (size_t)N);
            dumpbuf("vel", b_vel, 16 * (size_t)N);
        }
        glFinish();
        double t5 = now_s();
        check_gl("step");
        fprintf(tf, "%d %.6f %.6f %.6f %.6f %.6f %.6f\n", step, t1 - t0, t2 - t1, t3 - t2, t4 - t3, t5 - t4, t5 - t0);
        fflush(tf);
        if (step >= job.warmup) {
            double d[6] = {t1 - t0, t2 - t1, t3 - t2, t4 - t3, t5 - t4, t5 - t0};
            for (int k = 0; k < 6; k++) acc[k] += d[k];
            timed++;
        }
    }
    if (timed) {
        fprintf(tf, "# mean over %d steps: predict %.6f sort %.6f cells %.6f solver %.6f update_vorticity %.6f total %.6f\n",
                timed, acc[0] / timed, acc[1] / timed, acc[2] / timed, acc[3] / timed, acc[4] / timed, acc[5] / timed);
    }
    fclose(tf);
    write_file(outdir + "/final_pos.bin", read_buffer(b_pos, 16 * (size_t)N).data(), 16 * (size_t)N);
    write_file(outdir + "/final_vel.bin", read_buffer(b_vel, 16 * (size_t)N).data(), 16 * (size_t)N);
    write_file(outdir + "/final_vort.bin", read_buffer(b_vort, 4 * (size_t)N).data(), 4 * (size_t)N);
    write_file(outdir + "/final_end.bin", read_tex3d(t_end, C).data(), 4 * C);
    write_file(outdir + "/final_sorted.bin", read_buffer(rs_buffer, 16 * (size_t)N).data(), 16 * (size_t)N);
    write_file(outdir + "/final_highlight.bin", read_buffer(b_highlight, 4 * (size_t)N).data(), 4 * (size_t)N);
    check_gl("final readback");
    fprintf(stderr, "glsl_runner: %d steps done\n", job.nsteps);
    return 0;
}
