// Host restatement of the leaf filter's arithmetic -- the plane record of kBuildPrisms and the
// bound of leafLowerBound (meshlib_amd/csrc/mi_kernels.hip) -- and of the leaf test it guards (the
// FP64 closest point of closestPointInTriangle, source/MRMesh/MRClosestPointInTriangle.h:27-87,
// rounded to float and squared in float).  The filter is exact only if the bound never exceeds the
// leaf test's result; this checks that claim on random and adversarial triangles (slivers, needles,
// collinear, scales 1e-4..1e4, offsets up to 1e5) and points on, near and far from them.  The GPU
// kernels' side of the same claim is tests/test_gpu_leaf_filter.py.
// usage: leaf_bound_fuzz SEED COUNT  -> exit 0 when no sample violates the bound
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <random>
#include <algorithm>
#include <cstdlib>
static float roundUpF(double x) { float f = (float)x; if ((double)f < x) f = std::nextafter(f, INFINITY); return f; }
static float roundDownF(double x) { float f = (float)x; if ((double)f > x) f = std::nextafter(f, -INFINITY); return f; }
struct Rec { float r[20]; };
static Rec build(const float* a, const float* b, const float* c) {
    Rec R{}; float* out = R.r;
    const double v[3][3] = {{a[0],a[1],a[2]},{b[0],b[1],b[2]},{c[0],c[1],c[2]}};
    double r = 0; for (int j=0;j<3;++j) r = std::fmax(r, std::fabs(v[j][0])+std::fabs(v[j][1])+std::fabs(v[j][2]));
    const double e0[3] = {v[1][0]-v[0][0], v[1][1]-v[0][1], v[1][2]-v[0][2]};
    const double e1[3] = {v[2][0]-v[0][0], v[2][1]-v[0][1], v[2][2]-v[0][2]};
    const double nn[3] = {e0[1]*e1[2]-e0[2]*e1[1], e0[2]*e1[0]-e0[0]*e1[2], e0[0]*e1[1]-e0[1]*e1[0]};
    const double len = std::sqrt(nn[0]*nn[0]+nn[1]*nn[1]+nn[2]*nn[2]);
    if (!(r < INFINITY) || !(len > 0) || !(len < INFINITY)) return R;
    const double nd[3] = {nn[0]/len, nn[1]/len, nn[2]/len};
    float pl[4][3]; for (int k=0;k<3;++k) pl[0][k] = (float)nd[k];
    for (int e=0;e<3;++e) {
        const double* p0 = v[e]; const double* p1 = v[(e+1)%3];
        const double d[3] = {p1[0]-p0[0], p1[1]-p0[1], p1[2]-p0[2]};
        double mm[3] = {d[1]*nd[2]-d[2]*nd[1], d[2]*nd[0]-d[0]*nd[2], d[0]*nd[1]-d[1]*nd[0]};
        const double t = mm[0]*nd[0]+mm[1]*nd[1]+mm[2]*nd[2];
        for (int k=0;k<3;++k) mm[k] -= t*nd[k];
        const double ml = std::sqrt(mm[0]*mm[0]+mm[1]*mm[1]+mm[2]*mm[2]);
        for (int k=0;k<3;++k) pl[1+e][k] = ml > 0 ? (float)(mm[k]/ml) : 0.f;
    }
    double hi[4], lo = INFINITY; for (int q=0;q<4;++q) hi[q] = -INFINITY;
    for (int j=0;j<3;++j) for (int q=0;q<4;++q) {
        const double s = (double)pl[q][0]*v[j][0] + (double)pl[q][1]*v[j][1] + (double)pl[q][2]*v[j][2];
        hi[q] = std::fmax(hi[q], s); if (q==0) lo = std::fmin(lo, s);
    }
    const double marg = r * 0x1p-40;
    float o[20] = {pl[0][0],pl[0][1],pl[0][2],roundDownF(lo-marg), roundUpF(hi[0]+marg),pl[1][0],pl[1][1],pl[1][2],
                   roundUpF(hi[1]+marg),pl[2][0],pl[2][1],pl[2][2], roundUpF(hi[2]+marg),pl[3][0],pl[3][1],pl[3][2],
                   roundUpF(hi[3]+marg), roundUpF(r), 0.f, 0.f};
    for (int k=0;k<20;++k) out[k]=o[k];
    return R;
}
static float lbound(const Rec& R, float px, float py, float pz) {
    const float* q = R.r;
    const float np = std::fmaf(q[2], pz, std::fmaf(q[1], py, q[0]*px));
    const float a = std::fmax(np - q[4], q[3] - np);
    const float e0 = std::fmaf(q[7], pz, std::fmaf(q[6], py, std::fmaf(q[5], px, -q[8])));
    const float e1 = std::fmaf(q[11], pz, std::fmaf(q[10], py, std::fmaf(q[9], px, -q[12])));
    const float e2 = std::fmaf(q[15], pz, std::fmaf(q[14], py, std::fmaf(q[13], px, -q[16])));
    const float b = std::fmax(e0, std::fmax(e1, e2));
    const float eps = (std::fabs(px)+std::fabs(py)+std::fabs(pz)+q[17]) * 0x1p-20f;
    const float as = std::fmax(a-eps, 0.f), bs = std::fmax(b-eps, 0.f);
    return std::fmaf(as, as, bs*bs) * (1.f - 0x1p-13f);
}
static double dot3(double ax,double ay,double az,double bx,double by,double bz){return ax*bx+ay*by+az*bz;}
static void cp(const double p[3], const double a[3], const double b[3], const double c[3], double pr[3]) {
    const double abx=b[0]-a[0],aby=b[1]-a[1],abz=b[2]-a[2], acx=c[0]-a[0],acy=c[1]-a[1],acz=c[2]-a[2];
    const double apx=p[0]-a[0],apy=p[1]-a[1],apz=p[2]-a[2];
    const double d1=dot3(abx,aby,abz,apx,apy,apz), d2=dot3(acx,acy,acz,apx,apy,apz);
    if (d1<=0&&d2<=0){pr[0]=a[0];pr[1]=a[1];pr[2]=a[2];return;}
    const double bpx=p[0]-b[0],bpy=p[1]-b[1],bpz=p[2]-b[2];
    const double d3=dot3(abx,aby,abz,bpx,bpy,bpz), d4=dot3(acx,acy,acz,bpx,bpy,bpz);
    if (d3>=0&&d4<=d3){pr[0]=b[0];pr[1]=b[1];pr[2]=b[2];return;}
    const double cpx=p[0]-c[0],cpy=p[1]-c[1],cpz=p[2]-c[2];
    const double d5=dot3(abx,aby,abz,cpx,cpy,cpz), d6=dot3(acx,acy,acz,cpx,cpy,cpz);
    if (d6>=0&&d5<=d6){pr[0]=c[0];pr[1]=c[1];pr[2]=c[2];return;}
    const double vc=d1*d4-d3*d2;
    if (vc<=0&&d1>=0&&d3<=0){const double v=d1/(d1-d3);pr[0]=a[0]+v*abx;pr[1]=a[1]+v*aby;pr[2]=a[2]+v*abz;return;}
    const double vb=d5*d2-d1*d6;
    if (vb<=0&&d6<=0&&d2>=0){const double v=d2/(d2-d6);pr[0]=a[0]+v*acx;pr[1]=a[1]+v*acy;pr[2]=a[2]+v*acz;return;}
    const double va=d3*d6-d5*d4;
    if (va<=0){ if(d4<d3){pr[0]=b[0];pr[1]=b[1];pr[2]=b[2];return;} if(d5<d6){pr[0]=c[0];pr[1]=c[1];pr[2]=c[2];return;}
        const double v=(d4-d3)/((d4-d3)+(d5-d6)); pr[0]=b[0]+v*(c[0]-b[0]);pr[1]=b[1]+v*(c[1]-b[1]);pr[2]=b[2]+v*(c[2]-b[2]);return;}
    const double den=1/(va+vb+vc), v=vb*den, w=vc*den;
    pr[0]=a[0]+v*abx+w*acx; pr[1]=a[1]+v*aby+w*acy; pr[2]=a[2]+v*abz+w*acz;
}
static float leaf(const float* a,const float* b,const float* c,float px,float py,float pz){
    const double p[3]={px,py,pz},A[3]={a[0],a[1],a[2]},B[3]={b[0],b[1],b[2]},C[3]={c[0],c[1],c[2]}; double pr[3];
    cp(p,A,B,C,pr); const float dx=(float)pr[0]-px, dy=(float)pr[1]-py, dz=(float)pr[2]-pz; return dx*dx+dy*dy+dz*dz;
}
int main(int argc, char** argv) {
    std::mt19937_64 g(argc > 1 ? atoll(argv[1]) : 1);
    const long count = argc > 2 ? atol(argv[2]) : 4000000;
    std::uniform_real_distribution<double> U(-1, 1);
    long n = 0, bad = 0, tight = 0, kept = 0; double worst = 0;
    for (long it = 0; it < count; ++it) {
        const int kind = it % 8;
        double scale = std::pow(10.0, (int)(g() % 9) - 4);        // 1e-4 .. 1e4
        double off = (g() % 3 == 0) ? std::pow(10.0, (int)(g() % 7) - 1) : 0;  // offset 0 or 0.1..1e5
        float a[3], b[3], c[3];
        for (int k=0;k<3;++k){ a[k]=(float)(off+scale*U(g)); b[k]=(float)(off+scale*U(g)); c[k]=(float)(off+scale*U(g)); }
        if (kind == 1) for (int k=0;k<3;++k) c[k] = (float)(a[k] + (b[k]-a[k])*0.5 + scale*1e-6*U(g));     // sliver
        if (kind == 2) for (int k=0;k<3;++k) b[k] = (float)(a[k] + scale*1e-6*U(g));                      // needle
        if (kind == 3) for (int k=0;k<3;++k) c[k] = (float)(a[k] + (b[k]-a[k])*0.3);                      // collinear
        Rec R = build(a,b,c);
        for (int t = 0; t < 8; ++t) {
            float p[3];
            const int pk = t % 4;
            double u = std::fabs(U(g)), w = std::fabs(U(g)); if (u + w > 1) { u = 1-u; w = 1-w; }
            double h = scale * std::pow(10.0, -(int)(g() % 9)) * U(g);
            for (int k=0;k<3;++k) {
                double s = a[k] + u*(b[k]-a[k]) + w*(c[k]-a[k]);
                if (pk == 0) p[k] = (float)(s + h*U(g));                 // near the surface
                else if (pk == 1) p[k] = (float)((k==0?a:k==1?b:c)[k] + h*U(g));
                else if (pk == 2) p[k] = (float)(a[k] + 0.5*(b[k]-a[k]) + h*U(g));
                else p[k] = (float)(off + 4*scale*U(g));
            }
            if (pk == 1) { p[0]=(float)(a[0]+h*U(g)); p[1]=(float)(a[1]+h*U(g)); p[2]=(float)(a[2]+h*U(g)); }
            const float lb = lbound(R, p[0], p[1], p[2]);
            const float d = leaf(a,b,c,p[0],p[1],p[2]);
            ++n;
            if (lb > d) { ++bad; if (bad < 10) printf("BAD kind %d pk %d lb %.9g d %.9g\n", kind, pk, lb, d); }
            if (d > 0) { worst = std::max(worst, (double)lb/d); if (lb >= 0.5*d) ++tight; }
            if (!(lb > 0)) ++kept;
        }
    }
    printf("samples %ld bad %ld  max lb/d %.9f  lb>=d/2 %.3f  lb==0 %.3f\n", n, bad, worst, (double)tight/n, (double)kept/n);
    return bad != 0;
}
