classdef GpuApplicator
% GpuApplicator: The core product of a GPU product, the counterpart of
%   rooftopMoM.assembler.greensMatrix.Applicator (Applicator.m:1-48): apply, memory, workingMemory and spectra on a
%   librooftop product handle. It skips the checks of Product.apply, so the reference's solvers iterate with it
%   (Product.m:41-45). The owning rooftopMoM.gpu product frees the handle.

    properties (Access = private)
        handle_ uint64;
        precision_ double;
    end

    methods
        function obj = GpuApplicator(handle, precision)
            obj.handle_ = handle;
            obj.precision_ = precision;
        end

        function output = apply(obj, input)
        % apply: Returns the product with the Nx1 (or NxR) vectors of rooftop coefficients input.

            output = rooftop_mex('apply', obj.handle_, input, obj.precision_);
        end

        function bytes = memory(obj)
        % memory: Returns the device bytes of the stored spectra or coefficients.

            bytes = rooftop_mex('memory', obj.handle_);
        end

        function bytes = workingMemory(obj)
        % workingMemory: Returns the device bytes one application uses for one right-hand side.

            [~, bytes] = rooftop_mex('memory', obj.handle_);
        end

        function handle = handle(obj)
            handle = obj.handle_;
        end
    end
end
