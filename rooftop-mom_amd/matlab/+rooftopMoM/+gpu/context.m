function handle = context(device, budget)
% context: Returns the uint64 handle of the GPU context (device, HIP stream, device-byte ceiling) that the GPU products
%   and solves of this MATLAB session share, creating it on the first call (librooftop rt_ctx_create; the device
%   counterpart of rooftopMoM.Resources, src/+rooftopMoM/Resources.m:9-47).
%
%   Optional inputs (first call only):
%       device: HIP device index (default 0).
%       budget: Device bytes the context may hold, 0 for the device's free memory (default 0).

    persistent current;

    if isempty(current)
        if nargin < 1
            device = 0;
        end

        if nargin < 2
            budget = 0;
        end

        current = rooftop_mex('context', device, budget);
    end

    handle = current;
end
