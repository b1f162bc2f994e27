// Checks the MFMA operand/result lane layouts this library assumes (see
// momms_amd/csrc/mfma.hpp) against a CPU matmul; exit status = failures.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <vector>
typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v16f __attribute__((ext_vector_type(16)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
#define CK(x) do{hipError_t e=(x); if(e!=hipSuccess){printf("HIP error %s at %d\n",hipGetErrorString(e),__LINE__); return 1;}}while(0)

// f64 16x16x4: in[64] a, in[64..127] b ; out[64*4]
__global__ void kf64(const double* in, double* out){int l=threadIdx.x; v4d c={0,0,0,0};
 c=__builtin_amdgcn_mfma_f64_16x16x4f64(in[l],in[64+l],c,0,0,0); for(int r=0;r<4;r++) out[l*4+r]=c[r];}
// bf16 32x32x16: a[64][8], b[64][8]; out[64*16]
__global__ void kbf32(const v8bf* a,const v8bf* b,float* out){int l=threadIdx.x; v16f c={};
 c=__builtin_amdgcn_mfma_f32_32x32x16_bf16(a[l],b[l],c,0,0,0); for(int r=0;r<16;r++) out[l*16+r]=c[r];}
__global__ void kbf16(const v8bf* a,const v8bf* b,float* out){int l=threadIdx.x; v4f c={};
 c=__builtin_amdgcn_mfma_f32_16x16x32_bf16(a[l],b[l],c,0,0,0); for(int r=0;r<4;r++) out[l*4+r]=c[r];}
__global__ void kf32(const float* in,float* out){int l=threadIdx.x; v16f c={};
 c=__builtin_amdgcn_mfma_f32_32x32x2f32(in[l],in[64+l],c,0,0,0); for(int r=0;r<16;r++) out[l*16+r]=c[r];}
__global__ void ki8(const v4i* a,const v4i* b,int* out){int l=threadIdx.x; v16i c={};
 c=__builtin_amdgcn_mfma_i32_32x32x32_i8(a[l],b[l],c,0,0,0); for(int r=0;r<16;r++) out[l*16+r]=c[r];}

static uint16_t f2bf(float f){uint32_t u; memcpy(&u,&f,4); return (uint16_t)(u>>16);} // exact for small ints
int main(){
  int fails=0;
  { // f64 16x16x4
    double A[16][4],B[4][16]; for(int i=0;i<16;i++)for(int k=0;k<4;k++)A[i][k]=(i*7+k*3)%11-5;
    for(int k=0;k<4;k++)for(int j=0;j<16;j++)B[k][j]=(k*5+j*2)%13-6;
    double h[128]; for(int l=0;l<64;l++){h[l]=A[l%16][l/16]; h[64+l]=B[l/16][l%16];}
    double *din,*dout; CK(hipMalloc(&din,128*8)); CK(hipMalloc(&dout,256*8)); CK(hipMemcpy(din,h,128*8,hipMemcpyHostToDevice));
    kf64<<<1,64>>>(din,dout); CK(hipDeviceSynchronize()); double o[256]; CK(hipMemcpy(o,dout,256*8,hipMemcpyDeviceToHost));
    int bad=0; for(int l=0;l<64;l++)for(int r=0;r<4;r++){int i=l/16+4*r,j=l%16; double s=0; for(int k=0;k<4;k++)s+=A[i][k]*B[k][j]; if(s!=o[l*4+r])bad++;}
    printf("f64_16x16x4: %s (%d bad)\n",bad?"FAIL":"PASS",bad); fails+=bad!=0;
  }
  { // bf16 32x32x16
    float A[32][16],B[16][32]; for(int i=0;i<32;i++)for(int k=0;k<16;k++)A[i][k]=(i*7+k*3)%11-5;
    for(int k=0;k<16;k++)for(int j=0;j<32;j++)B[k][j]=(k*5+j*2)%13-6;
    uint16_t ha[64*8],hb[64*8]; for(int l=0;l<64;l++)for(int e=0;e<8;e++){ha[l*8+e]=f2bf(A[l%32][8*(l/32)+e]); hb[l*8+e]=f2bf(B[8*(l/32)+e][l%32]);}
    void *da,*db; float* dout; CK(hipMalloc(&da,1024)); CK(hipMalloc(&db,1024)); CK(hipMalloc(&dout,1024*4));
    CK(hipMemcpy(da,ha,1024,hipMemcpyHostToDevice)); CK(hipMemcpy(db,hb,1024,hipMemcpyHostToDevice));
    kbf32<<<1,64>>>((const v8bf*)da,(const v8bf*)db,dout); CK(hipDeviceSynchronize()); float o[1024]; CK(hipMemcpy(o,dout,4096,hipMemcpyDeviceToHost));
    int bad=0; for(int l=0;l<64;l++)for(int r=0;r<16;r++){int i=8*(r/4)+4*(l/32)+(r%4),j=l%32; float s=0; for(int k=0;k<16;k++)s+=A[i][k]*B[k][j]; if(s!=o[l*16+r])bad++;}
    printf("bf16_32x32x16: %s (%d bad)\n",bad?"FAIL":"PASS",bad); fails+=bad!=0;
    // 16x16x32
    float A2[16][32],B2[32][16]; for(int i=0;i<16;i++)for(int k=0;k<32;k++)A2[i][k]=(i*7+k*3)%11-5;
    for(int k=0;k<32;k++)for(int j=0;j<16;j++)B2[k][j]=(k*5+j*2)%13-6;
    for(int l=0;l<64;l++)for(int e=0;e<8;e++){ha[l*8+e]=f2bf(A2[l%16][8*(l/16)+e]); hb[l*8+e]=f2bf(B2[8*(l/16)+e][l%16]);}
    CK(hipMemcpy(da,ha,1024,hipMemcpyHostToDevice)); CK(hipMemcpy(db,hb,1024,hipMemcpyHostToDevice));
    kbf16<<<1,64>>>((const v8bf*)da,(const v8bf*)db,dout); CK(hipDeviceSynchronize()); CK(hipMemcpy(o,dout,1024,hipMemcpyDeviceToHost));
    bad=0; for(int l=0;l<64;l++)for(int r=0;r<4;r++){int i=4*(l/16)+r,j=l%16; float s=0; for(int k=0;k<32;k++)s+=A2[i][k]*B2[k][j]; if(s!=o[l*4+r])bad++;}
    printf("bf16_16x16x32: %s (%d bad)\n",bad?"FAIL":"PASS",bad); fails+=bad!=0;
  }
  { // f32 32x32x2
    float A[32][2],B[2][32]; for(int i=0;i<32;i++)for(int k=0;k<2;k++)A[i][k]=(i*7+k*3)%11-5;
    for(int k=0;k<2;k++)for(int j=0;j<32;j++)B[k][j]=(k*5+j*2)%13-6;
    float h[128]; for(int l=0;l<64;l++){h[l]=A[l%32][l/32]; h[64+l]=B[l/32][l%32];}
    float *din,*dout; CK(hipMalloc(&din,512)); CK(hipMalloc(&dout,4096)); CK(hipMemcpy(din,h,512,hipMemcpyHostToDevice));
    kf32<<<1,64>>>(din,dout); CK(hipDeviceSynchronize()); float o[1024]; CK(hipMemcpy(o,dout,4096,hipMemcpyDeviceToHost));
    int bad=0; for(int l=0;l<64;l++)for(int r=0;r<16;r++){int i=8*(r/4)+4*(l/32)+(r%4),j=l%32; float s=0; for(int k=0;k<2;k++)s+=A[i][k]*B[k][j]; if(s!=o[l*16+r])bad++;}
    printf("f32_32x32x2: %s (%d bad)\n",bad?"FAIL":"PASS",bad); fails+=bad!=0;
  }
  { // i8 32x32x32
    int8_t A[32][32],B[32][32]; for(int i=0;i<32;i++)for(int k=0;k<32;k++)A[i][k]=(i*7+k*3)%23-11;
    for(int k=0;k<32;k++)for(int j=0;j<32;j++)B[k][j]=(k*5+j*2)%29-14;
    int8_t ha[1024],hb[1024]; for(int l=0;l<64;l++)for(int e=0;e<16;e++){ha[l*16+e]=A[l%32][16*(l/32)+e]; hb[l*16+e]=B[16*(l/32)+e][l%32];}
    void *da,*db; int* dout; CK(hipMalloc(&da,1024)); CK(hipMalloc(&db,1024)); CK(hipMalloc(&dout,4096));
    CK(hipMemcpy(da,ha,1024,hipMemcpyHostToDevice)); CK(hipMemcpy(db,hb,1024,hipMemcpyHostToDevice));
    ki8<<<1,64>>>((const v4i*)da,(const v4i*)db,dout); CK(hipDeviceSynchronize()); int o[1024]; CK(hipMemcpy(o,dout,4096,hipMemcpyDeviceToHost));
    int bad=0; for(int l=0;l<64;l++)for(int r=0;r<16;r++){int i=8*(r/4)+4*(l/32)+(r%4),j=l%32; int s=0; for(int k=0;k<32;k++)s+=A[i][k]*B[k][j]; if(s!=o[l*16+r])bad++;}
    printf("i8_32x32x32: %s (%d bad)\n",bad?"FAIL":"PASS",bad); fails+=bad!=0;
  }
  hipDeviceProp_t p; CK(hipGetDeviceProperties(&p,0));
  printf("device %s CUs=%d clock=%d kHz lds/blk=%zu l2=%d regs/blk=%d\n",p.gcnArchName,p.multiProcessorCount,p.clockRate,p.sharedMemPerBlock,p.l2CacheSize,p.regsPerBlock);
  return fails;
}
