// Type declarations of fluid-render.js, matching render/fluidRender.ts:10-488.
export interface CubemapFaces {
    size: number;          // texels per face edge
    data: Float32Array;    // 6 faces (+X, -X, +Y, -Y, +Z, -Z) x size^2 RGBA, rows top to bottom
}

export interface PosVelSource {
    posvelDevicePtr(): bigint;
    synchronize?(): void;
}

export type ImageName = 'depth' | 'filteredDepth' | 'thickness' | 'filteredThickness' | 'color';

export declare class FluidRenderer {
    width: number;
    height: number;
    radius: number;
    fov: number;
    constructor(device: unknown, canvas: { width: number; height: number },
                presentationFormat: unknown, radius: number, fov: number,
                posvelBuffer: PosVelSource | bigint | ArrayBuffer | ArrayBufferView,
                renderUniformBuffer: ArrayBuffer, cubemapTextureView: CubemapFaces);
    execute(context: unknown, commandEncoder: unknown, numParticles: number,
            sphereRenderFl: boolean): void;
    readImage(name: ImageName, dst?: Float32Array): Float32Array;
    readColor(dst?: Float32Array): Float32Array;
    colorDevicePtr(): bigint;
    timeFrames(numParticles: number, sphereRenderFl: boolean, frames: number): number;
    destroy(): void;
}
