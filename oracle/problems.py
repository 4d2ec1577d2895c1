"""Raw problem definitions used by the oracle tests.

REF_DEFAULT restates the constants hard-coded in cnn/intensity/abbe/intensity.cpp:44-87,
cnn/intensity/m3d/m3d.cpp:44-111 and the material table of cnn/include/header.h:1217-1230.
The C1/C2 problems follow the workload descriptions of the benchmark configs
(1-GPU, FP64).
"""

REF_STACK = dict(
    n_mo=0.9238 + 0.006435j, n_si=0.999 + 0.001826j, n_ru=0.8863 + 0.01706j,
    n_mosi2=0.9693 + 0.004333j, n_rusi=0.9099 + 0.01547j, n_sio2=0.978 + 0.01083j,
    d_mo=2.052, d_mosi=1.661, d_si=2.283, d_simo=1.045, d_siru=0.8, d_ru=2.5, NML=40,
)

# 40 x Mo/Si (2.8/4.2 nm), Ru cap 2.5 nm, interlayers of zero thickness, SiO2 substrate.
HNA_STACK = dict(REF_STACK, n_mo=0.924 + 0.0064j, n_si=0.999 + 0.0018j,
                 d_mo=2.8, d_si=4.2, d_mosi=0.0, d_simo=0.0, d_siru=0.0, d_ru=2.5)

REF_DEFAULT = dict(
    NDIVX=2048, NDIVY=2048, MX=4, MY=4, NA=0.33, wavelength=13.5, theta0=-6.0, azimuth=0.0,
    eabs=[(0.9567 + 0.0343j) ** 2], dabs=[60.0], z0=60.0 + 42.0, z=0.0, stack=REF_STACK,
    source=dict(type=1, sigma1=0.9, sigma2=0.55, openangle=90.0),
)


def hna(ndiv, K, source, ndivs=None):
    """High-NA anamorphic problem (NA 0.55, 4x/8x, CRAO 5.355 deg, 60 nm absorber)."""
    return dict(
        NDIVX=ndiv, NDIVY=ndiv, MX=4, MY=8, NA=0.55, wavelength=13.5, theta0=-5.355, azimuth=0.0,
        eabs=[(0.947 + 0.031j) ** 2], dabs=[60.0], z0=60.0 + 42.0, z=0.0, stack=HNA_STACK,
        orders=("rect", K), source=source, ndivs=ndivs,
    )


def layer_list(st):
    """Bottom-to-top (eps, thickness) list of the multilayer cascade.

    Order of header.h:294-362: substrate SiO2, Mo, then per pair
    MoSi2(d_mosi), Si, [SiMo interlayer (MoSi2 material, d_simo), Mo] ..., the last
    Si is followed by RuSi(d_siru) and the Ru cap(d_ru)."""
    e = {k[2:]: v * v for k, v in st.items() if k.startswith("n_")}
    L = [(e["sio2"], 0.0), (e["mo"], st["d_mo"])]
    NML = st["NML"]
    for i in range(NML - 1, -1, -1):
        if i < NML - 1:
            L.append((e["mo"], st["d_mo"]))
        L.append((e["mosi2"], st["d_mosi"]))
        L.append((e["si"], st["d_si"]))
        if i > 0:
            L.append((e["mosi2"], st["d_simo"]))
        else:
            L.append((e["rusi"], st["d_siru"]))
    L.append((e["ru"], st["d_ru"]))
    return L
